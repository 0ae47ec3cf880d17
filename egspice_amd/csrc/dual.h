// Forward-mode dual numbers over a device's local unknowns (the reference's
// `Dual` of src/device/eval.zig:107-350, shared by host and device). The
// lanes follow the reference's formulas: the product rule and the
// quotient's numerator fused into one rounding (its `mulAddV` is a fused
// multiply-add on every FMA target, eval.zig:56-61, 346-348: x86-64 with
// FMA and amdgcn), a quotient's lanes scaled by the reciprocal, and the
// sqrt/pow slopes with their zero and non-finite guards (:260-316).
#pragma once

#include "emath.h"

namespace esp {

constexpr int DW = 4;

struct Dual {
    double v;
    double d[DW];
};

ESP_HD Dual dc(double v)
{
    Dual r;
    r.v = v;
    for (int i = 0; i < DW; i++) r.d[i] = 0.0;
    return r;
}
ESP_HD Dual dvar(double v, int k)
{
    Dual r = dc(v);
    r.d[k] = 1.0;
    return r;
}
ESP_HD Dual dadd(Dual a, Dual b)
{
    Dual r;
    r.v = a.v + b.v;
    for (int i = 0; i < DW; i++) r.d[i] = a.d[i] + b.d[i];
    return r;
}
ESP_HD Dual dsub(Dual a, Dual b)
{
    Dual r;
    r.v = a.v - b.v;
    for (int i = 0; i < DW; i++) r.d[i] = a.d[i] - b.d[i];
    return r;
}
ESP_HD Dual dneg(Dual a)
{
    Dual r;
    r.v = -a.v;
    for (int i = 0; i < DW; i++) r.d[i] = -a.d[i];
    return r;
}
ESP_HD Dual dmul(Dual a, Dual b)
{
    Dual r;
    r.v = a.v * b.v;
    for (int i = 0; i < DW; i++) r.d[i] = __builtin_fma(b.d[i], a.v, a.d[i] * b.v);
    return r;
}
ESP_HD Dual ddiv(Dual a, Dual b)
{
    Dual r;
    const double inv = 1.0 / b.v;
    r.v = a.v / b.v;
    for (int i = 0; i < DW; i++) r.d[i] = __builtin_fma(b.d[i], -r.v, a.d[i]) * inv;
    return r;
}
ESP_HD Dual dscale(double s, Dual a)
{
    Dual r;
    r.v = s * a.v;
    for (int i = 0; i < DW; i++) r.d[i] = s * a.d[i];
    return r;
}
ESP_HD Dual ddivs(Dual a, double s)
{
    Dual r;
    const double inv = 1.0 / s;
    r.v = a.v / s;
    for (int i = 0; i < DW; i++) r.d[i] = a.d[i] * inv;
    return r;
}
ESP_HD Dual daddc(Dual a, double c)
{
    a.v = a.v + c;
    return a;
}
ESP_HD Dual dcsub(double c, Dual a)
{
    Dual r;
    r.v = c - a.v;
    for (int i = 0; i < DW; i++) r.d[i] = -a.d[i];
    return r;
}
ESP_HD Dual dexp(Dual a)
{
    Dual r;
    r.v = e_exp(a.v);
    for (int i = 0; i < DW; i++) r.d[i] = r.v * a.d[i];
    return r;
}
ESP_HD Dual dsqrt(Dual a)
{
    Dual r;
    r.v = __builtin_sqrt(a.v);
    const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
    for (int i = 0; i < DW; i++) r.d[i] = a.d[i] * h;
    return r;
}
ESP_HD Dual dpowc(Dual a, double c)
{
    Dual r;
    r.v = e_pow(a.v, c);
    double g = a.v != 0.0 ? c * r.v / a.v : c * e_pow(a.v, c - 1.0);
    if (!(fabs_(g) < inf_())) g = 0.0;
    for (int i = 0; i < DW; i++) r.d[i] = a.d[i] * g;
    return r;
}
ESP_HD Dual dminc(Dual a, double c) { return c < a.v ? dc(c) : a; }
ESP_HD Dual dmaxc(Dual a, double c) { return c > a.v ? dc(c) : a; }

}  // namespace esp
