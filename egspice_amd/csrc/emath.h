// exp/log/pow of the lane engine, identical on host and device.
//
// The reference evaluates every device with one set of transcendental
// bodies compiled for both the host and the GPU (the third-party Gompute
// math, src/device/eval.zig:38-43), so a lane is bitwise its host run.
// These are the engine's own bodies from textbook methods:
//   exp: Cody-Waite reduction with a two-word ln 2 and the degree-13 Taylor
//        polynomial of e^r by Horner, |r| <= ln2/2;
//   log: the atanh series of log(1+f), f = m - 1, m in [sqrt(1/2),
//        sqrt(2)), in the form f - (f^2/2 - s (f^2/2 + R)) of Sun's fdlibm;
//   pow: exp(y log x) with the IEEE special cases;
//   sin: fdlibm's pi/2 reduction and minimax kernels (below).
#pragma once

#include "common.h"

namespace esp {

constexpr double LN2_HI = 0.6931471803691238;
constexpr double LN2_LO = 1.9082149292705877e-10;
constexpr double INV_LN2 = 1.4426950408889634;

ESP_HD double pow2i(int k) { return from_bits((uint64_t)(k + 1023) << 52); }

ESP_HD double e_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return inf_();
    if (x < -745.1332191019412) return 0.0;
    double kd = __builtin_rint(x * INV_LN2);
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

ESP_HD double e_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -inf_();
    if (x == inf_()) return x;
    int k = 0;
    if (x < 2.2250738585072014e-308) {
        x = x * 18014398509481984.0;
        k = -54;
    }
    uint64_t b = bits(x);
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

ESP_HD double e_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (x != x || y != y) return __builtin_nan("");
    if (x == 0.0) return y > 0.0 ? 0.0 : inf_();
    if (x == inf_()) return y > 0.0 ? inf_() : 0.0;
    if (x < 0.0) {
        double yi = __builtin_rint(y);
        if (yi != y) return __builtin_nan("");
        double r = e_exp(y * e_log(-x));
        double half = y * 0.5;
        return __builtin_rint(half) != half ? -r : r;
    }
    return e_exp(y * e_log(x));
}

// sin, after Sun's fdlibm (s_sin.c, k_sin.c, k_cos.c, e_rem_pio2.c):
// x = n pi/2 + (y0 + y1) by Cody-Waite with pi/2 in three 33-bit pieces
// (valid for |x| < 2^20 pi/2, far beyond any source argument here), then
// the degree-13 odd / degree-14 even minimax polynomials on [-pi/4, pi/4].
constexpr double SIN_S[6] = {-1.66666666666666324348e-01, 8.33333333332248946124e-03, -1.98412698298579493134e-04,
                             2.75573137070700676789e-06,  -2.50507602534068634195e-08, 1.58969099521155010221e-10};
constexpr double COS_C[6] = {4.16666666666666019037e-02,  -1.38888888888741095749e-03, 2.48015872894767294178e-05,
                             -2.75573143513906633035e-07, 2.08757232129817482790e-09,  -1.13596475577881948265e-11};

ESP_HD double k_sin(double x, double y, int iy)
{
    const double z = x * x, v = z * x;
    const double r = SIN_S[1] + z * (SIN_S[2] + z * (SIN_S[3] + z * (SIN_S[4] + z * SIN_S[5])));
    if (iy == 0) return x + v * (SIN_S[0] + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * SIN_S[0]);
}

ESP_HD double k_cos(double x, double y)
{
    const uint32_t ix = (uint32_t)(bits(x) >> 32) & 0x7fffffffu;
    if (ix < 0x3e400000u && (int)x == 0) return 1.0;
    const double z = x * x;
    const double r = z * (COS_C[0] + z * (COS_C[1] + z * (COS_C[2] + z * (COS_C[3] + z * (COS_C[4] + z * COS_C[5])))));
    if (ix < 0x3FD33333u) return 1.0 - (0.5 * z - (z * r - x * y));
    const double qx = ix > 0x3fe90000u ? 0.28125 : from_bits((uint64_t)(ix - 0x00200000u) << 32);
    const double hz = 0.5 * z - qx, a = 1.0 - qx;
    return a - (hz - (z * r - x * y));
}

ESP_HD int rem_pio2(double x, double &y0, double &y1)
{
    const double invpio2 = 6.36619772367581382433e-01, pio2_1 = 1.57079632673412561417e+00,
                 pio2_1t = 6.07710050650619224932e-11, pio2_2 = 6.07710050630396597660e-11,
                 pio2_2t = 2.02226624879595063154e-21, pio2_3 = 2.02226624871116645580e-21,
                 pio2_3t = 8.47842766036889956997e-32;
    const double ax = x < 0 ? -x : x;
    const int j = (int)((bits(ax) >> 52) & 0x7ff);
    const int n = (int)(ax * invpio2 + 0.5);
    const double fn = (double)n;
    double r = ax - fn * pio2_1, w = fn * pio2_1t;
    y0 = r - w;
    int i = j - (int)((bits(y0) >> 52) & 0x7ff);
    if (i > 16) {
        double t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        y0 = r - w;
        i = j - (int)((bits(y0) >> 52) & 0x7ff);
        if (i > 49) {
            t = r;
            w = fn * pio2_3;
            r = t - w;
            w = fn * pio2_3t - ((t - r) - w);
            y0 = r - w;
        }
    }
    y1 = (r - y0) - w;
    if (x < 0) {
        y0 = -y0;
        y1 = -y1;
        return -n;
    }
    return n;
}

ESP_HD double e_sin(double x)
{
    const uint32_t ix = (uint32_t)(bits(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) return ix < 0x3e400000u ? x : k_sin(x, 0.0, 0);
    if (ix >= 0x7ff00000u) return x - x;
    double y0, y1;
    const int n = rem_pio2(x, y0, y1);
    switch (n & 3) {
    case 0: return k_sin(y0, y1, 1);
    case 1: return k_cos(y0, y1);
    case 2: return -k_sin(y0, y1, 1);
    default: return -k_cos(y0, y1);
    }
}

// Verilog-A min/max on reals: the second operand wins only when strictly
// smaller (larger).
ESP_HD double va_min(double a, double b) { return b < a ? b : a; }
ESP_HD double va_max(double a, double b) { return b > a ? b : a; }

}  // namespace esp
