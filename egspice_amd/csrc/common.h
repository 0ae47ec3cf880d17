// Shared host/device helpers of the lane engine.
//
// Every body that runs on a lane is written once as ESP_HD code, so the
// host tools (preparation of device constants, the lead tape, tests of the
// C ABI) and the gfx950 kernels execute the same operations. The library
// is built with -ffp-contract=off: each arithmetic expression here is one
// IEEE operation, and f64 division and sqrt are correctly rounded on both
// sides.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define ESP_HD __host__ __device__ inline
#define ESP_NONE 0xffffffffu

namespace esp {

ESP_HD uint64_t bits(double d) { return __builtin_bit_cast(uint64_t, d); }
ESP_HD double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }
ESP_HD double fabs_(double x) { return from_bits(bits(x) & 0x7fffffffffffffffull); }
// LLVM maxnum/minnum: a NaN operand yields the other one.
ESP_HD double fmax_(double a, double b) { return a != a ? b : (b != b ? a : (a > b ? a : b)); }
ESP_HD double fmin_(double a, double b) { return a != a ? b : (b != b ? a : (a < b ? a : b)); }
ESP_HD bool isfinite_(double x) { return fabs_(x) < __builtin_huge_val(); }
ESP_HD double inf_() { return __builtin_huge_val(); }

// A lane's view of a lane-major plane: element i of lane l of N lives at
// base[i * N + l]. On the host with N = 1 it is a plain array.
template <class T>
struct LV {
    T *p;
    size_t s;
    ESP_HD T &operator[](size_t i) const { return p[i * s]; }
    ESP_HD LV<T> at(size_t off) const { return LV<T>{p + off * s, s}; }
};

template <class T>
ESP_HD LV<T> lane_view(T *base, size_t n_lanes, size_t lane) { return LV<T>{base + lane, n_lanes}; }

}  // namespace esp
