/* Transcendental bodies every device evaluation uses, host side of the
 * oracle. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference evaluates devices with the third-party Gompute library's
 * exp/log/pow, compiled once for the host and once for the GPU, so a lane
 * on the device is bitwise its host run (src/device/eval.zig:38-43,
 * 275-325). Gompute (OmarSiwy/Gompute#15ab5813, gompute-1.5.0) is not in
 * the reference tree, so its bodies cannot be restated. This file defines
 * the engine's own bodies from textbook methods, and the HIP engine carries
 * the same algorithm (egspice_amd/csrc/emath.h) so host and device agree bit
 * for bit, as the reference's do:
 *   exp: Cody-Waite reduction x = k ln2 + r with a two-word ln2, then the
 *        degree-13 Taylor polynomial of e^r (|r| <= ln2/2) by Horner.
 *   log: x = 2^k m, m in [sqrt(1/2), sqrt(2)), s = f/(2+f), f = m-1,
 *        log(1+f) = f - (f^2/2 - s (f^2/2 + R)), R = sum 2 z^j/(2j+1),
 *        z = s^2 (the atanh series), as in Sun's fdlibm e_log.c.
 *   pow: exp(y log x) for x > 0 with the IEEE special cases.
 *   sin: Sun's fdlibm s_sin.c: x = n pi/2 + (y0 + y1) by Cody-Waite with
 *        pi/2 in three 33-bit pieces (e_rem_pio2.c, |x| < 2^20 pi/2), then
 *        the k_sin.c / k_cos.c minimax polynomials on [-pi/4, pi/4].
 * Parity pinned by: tests/test_oracle_math.py (against the platform libm
 * within 2 ulp over 10^6 arguments, plus special values; sin within 1 ulp). Agreement with
 * Gompute's last bit: parity unpinned (library absent).
 * Compile with -ffp-contract=off: every operation here is one IEEE op.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static double from_bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t to_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

static const double LN2_HI = 0.6931471803691238;      /* 0x3FE62E42FEE00000 */
static const double LN2_LO = 1.9082149292705877e-10;
static const double INV_LN2 = 1.4426950408889634;

/* 2^k for k in [-1022, 1023] as a normal double. */
static double pow2i(int k) { return from_bits((uint64_t)(k + 1023) << 52); }

double om_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019412) return 0.0;
    double kd = rint(x * INV_LN2);
    int k = (int)kd;
    double hi = x - kd * LN2_HI;
    double r = hi - kd * LN2_LO;
    double p = 1.6059043836821613e-10;
    p = p * r + 2.08767569878681e-09;
    p = p * r + 2.505210838544172e-08;
    p = p * r + 2.755731922398589e-07;
    p = p * r + 2.7557319223985893e-06;
    p = p * r + 2.48015873015873e-05;
    p = p * r + 0.0001984126984126984;
    p = p * r + 0.001388888888888889;
    p = p * r + 0.008333333333333333;
    p = p * r + 0.041666666666666664;
    p = p * r + 0.16666666666666666;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    if (k > 1023) return p * pow2i(1023) * pow2i(k - 1023);
    if (k < -1022) return p * pow2i(k + 1000) * pow2i(-1000);
    return p * pow2i(k);
}

double om_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return x;
    int k = 0;
    if (x < 2.2250738585072014e-308) {
        x = x * 18014398509481984.0; /* 2^54 */
        k = -54;
    }
    uint64_t b = to_bits(x);
    k += (int)((b >> 52) & 0x7ff) - 1023;
    double m = from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        k += 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 0.09523809523809523;
    R = R * z + 0.10526315789473684;
    R = R * z + 0.11764705882352941;
    R = R * z + 0.13333333333333333;
    R = R * z + 0.15384615384615385;
    R = R * z + 0.18181818181818182;
    R = R * z + 0.2222222222222222;
    R = R * z + 0.2857142857142857;
    R = R * z + 0.4;
    R = R * z + 0.6666666666666666;
    R = R * z;
    double hfsq = 0.5 * f * f;
    double l1p = f - (hfsq - s * (hfsq + R));
    double kd = (double)k;
    return kd * LN2_HI + (l1p + kd * LN2_LO);
}

double om_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (x != x || y != y) return NAN;
    if (x == 0.0) return y > 0.0 ? 0.0 : INFINITY;
    if (x == INFINITY) return y > 0.0 ? INFINITY : 0.0;
    if (x < 0.0) {
        double yi = rint(y);
        if (yi != y) return NAN;
        double r = om_exp(y * om_log(-x));
        /* odd integer exponent keeps the sign */
        double half = y * 0.5;
        return rint(half) != half ? -r : r;
    }
    return om_exp(y * om_log(x));
}

/* Verilog-A min/max on reals: the second operand wins only when strictly
 * smaller (larger), so a NaN first operand propagates. */
double om_min(double a, double b) { return b < a ? b : a; }
double om_max(double a, double b) { return b > a ? b : a; }

static const double S_[6] = {-1.66666666666666324348e-01, 8.33333333332248946124e-03, -1.98412698298579493134e-04,
                             2.75573137070700676789e-06, -2.50507602534068634195e-08, 1.58969099521155010221e-10};
static const double C_[6] = {4.16666666666666019037e-02, -1.38888888888741095749e-03, 2.48015872894767294178e-05,
                             -2.75573143513906633035e-07, 2.08757232129817482790e-09, -1.13596475577881948265e-11};

static double kernel_sin(double x, double y, int iy)
{
    double z = x * x, v = z * x;
    double r = S_[1] + z * (S_[2] + z * (S_[3] + z * (S_[4] + z * S_[5])));
    if (iy == 0) return x + v * (S_[0] + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S_[0]);
}

static double kernel_cos(double x, double y)
{
    uint32_t ix = (uint32_t)(to_bits(x) >> 32) & 0x7fffffffu;
    if (ix < 0x3e400000u && (int)x == 0) return 1.0;
    double z = x * x;
    double r = z * (C_[0] + z * (C_[1] + z * (C_[2] + z * (C_[3] + z * (C_[4] + z * C_[5])))));
    if (ix < 0x3FD33333u) return 1.0 - (0.5 * z - (z * r - x * y));
    double qx = ix > 0x3fe90000u ? 0.28125 : from_bits((uint64_t)(ix - 0x00200000u) << 32);
    double hz = 0.5 * z - qx, a = 1.0 - qx;
    return a - (hz - (z * r - x * y));
}

static int rem_pio2(double x, double *y0, double *y1)
{
    const double invpio2 = 6.36619772367581382433e-01, pio2_1 = 1.57079632673412561417e+00,
                 pio2_1t = 6.07710050650619224932e-11, pio2_2 = 6.07710050630396597660e-11,
                 pio2_2t = 2.02226624879595063154e-21, pio2_3 = 2.02226624871116645580e-21,
                 pio2_3t = 8.47842766036889956997e-32;
    double ax = fabs(x);
    int j = (int)((to_bits(ax) >> 52) & 0x7ff);
    int n = (int)(ax * invpio2 + 0.5);
    double fn = (double)n;
    double r = ax - fn * pio2_1, w = fn * pio2_1t;
    *y0 = r - w;
    int i = j - (int)((to_bits(*y0) >> 52) & 0x7ff);
    if (i > 16) {
        double t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        *y0 = r - w;
        i = j - (int)((to_bits(*y0) >> 52) & 0x7ff);
        if (i > 49) {
            t = r;
            w = fn * pio2_3;
            r = t - w;
            w = fn * pio2_3t - ((t - r) - w);
            *y0 = r - w;
        }
    }
    *y1 = (r - *y0) - w;
    if (x < 0) {
        *y0 = -*y0;
        *y1 = -*y1;
        return -n;
    }
    return n;
}

double om_sin(double x)
{
    uint32_t ix = (uint32_t)(to_bits(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) return ix < 0x3e400000u ? x : kernel_sin(x, 0.0, 0);
    if (ix >= 0x7ff00000u) return x - x;
    double y0, y1;
    int n = rem_pio2(x, &y0, &y1);
    switch (n & 3) {
    case 0: return kernel_sin(y0, y1, 1);
    case 1: return kernel_cos(y0, y1);
    case 2: return -kernel_sin(y0, y1, 1);
    default: return -kernel_cos(y0, y1);
    }
}
