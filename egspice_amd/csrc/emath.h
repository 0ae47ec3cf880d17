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
//   pow: exp(y log x) with the IEEE special cases.
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

// Verilog-A min/max on reals: the second operand wins only when strictly
// smaller (larger).
ESP_HD double va_min(double a, double b) { return b < a ? b : a; }
ESP_HD double va_max(double a, double b) { return b > a ? b : a; }

}  // namespace esp
