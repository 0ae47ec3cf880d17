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
//   wyhash        the sites' name hash (std.hash.Wyhash as the reference
//                 calls it, src/frontend/netlist.zig:690-692): Wang Yi's
//                 published wyhash with its 64-bit multiply-fold mixer
//   name_key      src/frontend/netlist.zig:690-692
#pragma once

#include <cmath>
#include <cstddef>
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

// 128-bit product of a and b, folded: low ^ high.
inline void wy_mum(uint64_t &a, uint64_t &b)
{
    const unsigned __int128 r = (unsigned __int128)a * b;
    a = (uint64_t)r;
    b = (uint64_t)(r >> 64);
}
inline uint64_t wy_mix(uint64_t a, uint64_t b)
{
    wy_mum(a, b);
    return a ^ b;
}
inline uint64_t wy_read(const unsigned char *p, int k)
{
    uint64_t v = 0;
    for (int i = k - 1; i >= 0; i--) v = v << 8 | p[i];
    return v;
}

inline uint64_t wyhash(uint64_t seed, const void *key, size_t len)
{
    static const uint64_t secret[4] = {0xa0761d6478bd642full, 0xe7037ed1a0b428dbull, 0x8ebc6af09c88c6e3ull,
                                       0x589965cc75374cc3ull};
    const unsigned char *p = (const unsigned char *)key;
    seed ^= wy_mix(seed ^ secret[0], secret[1]);
    uint64_t a, b;
    if (len <= 16) {
        if (len >= 4) {
            const size_t q = (len >> 3) << 2;
            a = wy_read(p, 4) << 32 | wy_read(p + q, 4);
            b = wy_read(p + len - 4, 4) << 32 | wy_read(p + len - 4 - q, 4);
        } else if (len > 0) {
            a = (uint64_t)p[0] << 16 | (uint64_t)p[len >> 1] << 8 | p[len - 1];
            b = 0;
        } else {
            a = b = 0;
        }
    } else {
        size_t i = len;
        if (i > 48) {
            uint64_t s1 = seed, s2 = seed;
            do {
                seed = wy_mix(wy_read(p, 8) ^ secret[1], wy_read(p + 8, 8) ^ seed);
                s1 = wy_mix(wy_read(p + 16, 8) ^ secret[2], wy_read(p + 24, 8) ^ s1);
                s2 = wy_mix(wy_read(p + 32, 8) ^ secret[3], wy_read(p + 40, 8) ^ s2);
                p += 48;
                i -= 48;
            } while (i > 48);
            seed ^= s1 ^ s2;
        }
        while (i > 16) {
            seed = wy_mix(wy_read(p, 8) ^ secret[1], wy_read(p + 8, 8) ^ seed);
            i -= 16;
            p += 16;
        }
        a = wy_read(p + i - 16, 8);
        b = wy_read(p + i - 8, 8);
    }
    a ^= secret[1];
    b ^= seed;
    wy_mum(a, b);
    return wy_mix(a ^ secret[0] ^ len, b ^ secret[1]);
}

inline uint64_t name_key(uint64_t owner, const char *name, size_t len) { return site_key(owner, wyhash(0, name, len)); }

// The site owner of a distribution call in `.param name` text or on card
// `name` at the top level (netlist.zig:1832, 3570, 3826).
inline uint64_t top_owner(const char *name, size_t len) { return name_key(name_key(0, "", 0), name, len); }

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
