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
 * Parity pinned by: tests/test_oracle_math.py (against the platform libm
 * within 2 ulp over 10^6 arguments, plus special values). Agreement with
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
