/* Monte Carlo draws. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates the reference's counter-based sampler:
 *   or_site_key         src/frontend/netlist.zig:683-688 (splitmix64's
 *                       finalizer of owner ^ (salt + golden gamma))
 *   or_mc_uniform       src/frontend/variants.zig:1716-1720 (simple random
 *                       sampling: the top 53 bits, centred in their cell)
 *   or_normal_quantile  src/core/numerics.zig:416-432 (Acklam)
 *   or_trial_draw       src/frontend/variants.zig:1862-1895 (TrialDraw.sample)
 *   or_wyhash           the name hash of the sites (std.hash.Wyhash as
 *                       src/frontend/netlist.zig:690-692 calls it): Wang
 *                       Yi's published wyhash (64-bit multiply-fold mixer,
 *                       secret 0xa0761d6478bd642f, 0xe7037ed1a0b428db,
 *                       0x8ebc6af09c88c6e3, 0x589965cc75374cc3)
 *   or_param_site       a `.param`'s first distribution call's site,
 *                       siteKey(nameKey(nameKey(0, ""), name), 0)
 *                       (netlist.zig:3564-3601, levelPath :3666)
 * Parity pinned by: tests/test_mc.py (splitmix64's published output
 * sequence from state 0; the normal quantile against Python's
 * statistics.NormalDist within Acklam's 1.2e-9 bound and the reference's
 * own quantile tests, numerics.zig:520-545; TrialDraw's tests,
 * variants.zig:1917-1933); the wyhash keys by the reference's printed
 * Monte Carlo rows of docs/examples/monte_agauss.sp (docs/learn/
 * monte-carlo.md:62-69, 17 digits), which only the right hash reproduces.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

uint64_t or_site_key(uint64_t owner, uint64_t salt)
{
    uint64_t z = owner ^ (salt + 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

double or_mc_uniform(uint64_t seed, uint64_t site, uint32_t trial)
{
    const uint64_t bits = or_site_key(or_site_key(seed, trial), site);
    return ((double)(bits >> 11) + 0.5) * 0x1.0p-53;
}

double or_normal_quantile(double u)
{
    static const double a[] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                               1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                               6.680131188771972e+01, -1.328068155288572e+01};
    static const double c[] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                               -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
    static const double d[] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                               3.754408661907416e+00};
    const double low = 0.02425;
    if (u < low || u > 1 - low) {
        const double q = sqrt(-2 * log(u < low ? u : 1 - u));
        const double x = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
                         ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
        return u < low ? x : -x;
    }
    const double q = u - 0.5, r = q * q;
    return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
           (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1);
}

/* fn: 0 agauss, 1 gauss, 2 aunif, 3 unif, 4 limit. */
double or_trial_draw(uint64_t seed, uint32_t trial, uint64_t site, int fn, const double *args, uint32_t nargs)
{
    const double nom = args[0];
    const double width = nargs > 1 ? args[1] : 0;
    const int gaussian = fn == 0 || fn == 1;
    const double sigmas = gaussian && nargs > 2 && args[2] != 0 ? args[2] : 1;
    const uint32_t mult_at = gaussian ? 3 : 2;
    uint32_t mult = 1;
    if (fn != 4 && nargs > mult_at && args[mult_at] >= 1) mult = (uint32_t)(args[mult_at] < 1e6 ? args[mult_at] : 1e6);
    double scale;
    switch (fn) {
    case 0: scale = width / sigmas; break;
    case 1: scale = nom * width / sigmas; break;
    case 3: scale = nom * width; break;
    default: scale = width; break;
    }
    double best = 0;
    for (uint32_t m = 0; m < mult; m++) {
        const uint64_t key = or_site_key(site, m);
        const double u = or_mc_uniform(seed, key, trial);
        double dv;
        if (gaussian) dv = scale * or_normal_quantile(u);
        else if (fn == 4) dv = u < 0.5 ? -scale : scale;
        else dv = scale * (2 * u - 1);
        if (m == 0 || fabs(dv) > fabs(best)) best = dv;
    }
    return nom + best;
}

static void wy_mum(uint64_t *a, uint64_t *b)
{
    unsigned __int128 r = (unsigned __int128)*a * *b;
    *a = (uint64_t)r;
    *b = (uint64_t)(r >> 64);
}

static uint64_t wy_mix(uint64_t a, uint64_t b)
{
    wy_mum(&a, &b);
    return a ^ b;
}

static uint64_t wy_le(const unsigned char *p, int k)
{
    uint64_t v = 0;
    for (int i = 0; i < k; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

uint64_t or_wyhash(uint64_t seed, const char *key, size_t len)
{
    static const uint64_t S[4] = {0xa0761d6478bd642full, 0xe7037ed1a0b428dbull, 0x8ebc6af09c88c6e3ull,
                                  0x589965cc75374cc3ull};
    const unsigned char *p = (const unsigned char *)key;
    uint64_t a, b;
    seed ^= wy_mix(seed ^ S[0], S[1]);
    if (len <= 16) {
        if (len >= 4) {
            size_t q = (len >> 3) << 2;
            a = (wy_le(p, 4) << 32) | wy_le(p + q, 4);
            b = (wy_le(p + len - 4, 4) << 32) | wy_le(p + len - 4 - q, 4);
        } else if (len > 0) {
            a = ((uint64_t)p[0] << 16) | ((uint64_t)p[len >> 1] << 8) | p[len - 1];
            b = 0;
        } else {
            a = b = 0;
        }
    } else {
        size_t i = len;
        if (i > 48) {
            uint64_t s1 = seed, s2 = seed;
            do {
                seed = wy_mix(wy_le(p, 8) ^ S[1], wy_le(p + 8, 8) ^ seed);
                s1 = wy_mix(wy_le(p + 16, 8) ^ S[2], wy_le(p + 24, 8) ^ s1);
                s2 = wy_mix(wy_le(p + 32, 8) ^ S[3], wy_le(p + 40, 8) ^ s2);
                p += 48;
                i -= 48;
            } while (i > 48);
            seed ^= s1 ^ s2;
        }
        while (i > 16) {
            seed = wy_mix(wy_le(p, 8) ^ S[1], wy_le(p + 8, 8) ^ seed);
            i -= 16;
            p += 16;
        }
        a = wy_le(p + i - 16, 8);
        b = wy_le(p + i - 8, 8);
    }
    a ^= S[1];
    b ^= seed;
    wy_mum(&a, &b);
    return wy_mix(a ^ S[0] ^ (uint64_t)len, b ^ S[1]);
}

uint64_t or_param_site(const char *name, uint32_t ordinal)
{
    const uint64_t top = or_site_key(0, or_wyhash(0, "", 0));
    const uint64_t owner = or_site_key(top, or_wyhash(0, name, strlen(name)));
    return or_site_key(owner, ordinal);
}
