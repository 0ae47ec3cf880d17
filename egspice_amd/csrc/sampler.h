// Per-lane Monte Carlo draws (host side): the reference's counter-based
// sampler, so a trial's parameter values depend only on (seed, trial,
// site), never on the lane order or count.
//   site_key      src/frontend/netlist.zig:683-688 (splitmix64 finalizer)
//   mc_uniform    src/frontend/variants.zig:1716-1720 (simple random
//                 sampling: 53 bits into the open interval (0, 1))
//   normal_quantile  src/core/numerics.zig:416-432 (Acklam's rational
//                 approximation, relative error below 1.2e-9)
//   trial_draw    src/frontend/variants.zig:1862-1895 (agauss, gauss,
//                 aunif, unif, limit; with a multiplier m the draw of
//                 largest deviation among m)
#pragma once

#include <cmath>
#include <cstdint>

namespace esp {

enum { MC_AGAUSS = 0, MC_GAUSS = 1, MC_AUNIF = 2, MC_UNIF = 3, MC_LIMIT = 4 };

inline uint64_t site_key(uint64_t owner, uint64_t salt)
{
    uint64_t z = owner ^ (salt + 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

inline double mc_uniform(uint64_t seed, uint64_t site, uint32_t trial)
{
    const uint64_t bits = site_key(site_key(seed, trial), site);
    return ((double)(bits >> 11) + 0.5) * 0x1.0p-53;
}

// The coefficients are the published ones of Acklam's algorithm, as the
// reference writes them.
inline double normal_quantile(double u)
{
    static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                                1.383577518672690e+02,  -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                                6.680131188771972e+01, -1.328068155288572e+01};
    static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                                -2.549732539343734e+00, 4.374664141464968e+00,  2.938163982698783e+00};
    static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                                3.754408661907416e+00};
    const double low = 0.02425;
    if (u < low || u > 1 - low) {
        const double q = std::sqrt(-2 * std::log(u < low ? u : 1 - u));
        const double x = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
                         ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
        return u < low ? x : -x;
    }
    const double q = u - 0.5;
    const double r = q * q;
    return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
           (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1);
}

// One trial's draw of distribution `fn` at `site`; args = (nom, width
// [, sigmas][, mult]).
inline double trial_draw(uint64_t seed, uint32_t trial, uint64_t site, int fn, const double *args, uint32_t nargs)
{
    const double nom = args[0];
    const double width = nargs > 1 ? args[1] : 0;
    const bool gaussian = fn == MC_AGAUSS || fn == MC_GAUSS;
    const double sigmas = gaussian && nargs > 2 && args[2] != 0 ? args[2] : 1;
    const uint32_t mult_at = gaussian ? 3 : 2;
    uint32_t mult = 1;
    if (fn != MC_LIMIT && nargs > mult_at && args[mult_at] >= 1)
        mult = (uint32_t)(args[mult_at] < 1e6 ? args[mult_at] : 1e6);
    const double scale = fn == MC_AGAUSS ? width / sigmas
                       : fn == MC_GAUSS  ? nom * width / sigmas
                       : fn == MC_UNIF   ? nom * width
                                         : width;
    double best = 0;
    for (uint32_t m = 0; m < mult; m++) {
        const uint64_t key = site_key(site, m);
        const double u = mc_uniform(seed, key, trial);
        const double dv = gaussian ? scale * (1 * normal_quantile(u))
                        : fn == MC_LIMIT ? (u < 0.5 ? -scale : scale)
                                         : scale * (2 * u - 1);
        if (m == 0 || std::fabs(dv) > std::fabs(best)) best = dv;
    }
    return nom + best;
}

}  // namespace esp
