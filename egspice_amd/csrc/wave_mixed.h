// FP32 factor and triangular solves, recovered by FP64 iterative
// refinement: the lane's linear step of a Newton tick in mixed precision.
// The device evaluation, the residual, the matrix-vector products of the
// refinement and the Newton update stay FP64; only the LU (dense, partial
// pivoting by rows) and its two substitutions run in FP32. A refinement
// that does not reach FP64 accuracy hands the tick back to the FP64 factor.
//
// One wavefront owns the lane (n <= 64): thread c owns column c of the FP32
// factor for the factorisation's updates, thread r owns row r of the
// vectors, and the substitutions pass each solved component by a lane
// shuffle instead of through LDS.
#pragma once

#include "wave_lu.h"

namespace esp {

// LDS words (8 bytes) of the mixed step's arrays after the lane block: the
// FP32 matrix (n*n, column-major), and the pattern by rows (row pointers
// n+1, then (slot, column) pairs per entry) for the residual's rows.
ESP_HD size_t mixed_words(uint32_t n, uint32_t nnz)
{
    const size_t floats = (size_t)n * n;
    const size_t u32s = (size_t)n + 1 + 2 * (size_t)nnz;
    return (floats * 4 + u32s * 4 + 7) / 8;
}

struct MixedLU {
    float *F;        // n*n, column c at F + c*n
    uint32_t *rptr;  // n+1
    uint32_t *rent;  // 2 per entry: slot, column
};

ESP_HD MixedLU mixed_lu(double *at, uint32_t n)
{
    MixedLU m;
    m.F = reinterpret_cast<float *>(at);
    m.rptr = reinterpret_cast<uint32_t *>(m.F + (size_t)n * n);
    m.rent = m.rptr + n + 1;
    return m;
}

#if defined(__HIPCC__)

// The pattern by rows, once per launch (thread 0): each row's entries in
// ascending column order.
__device__ void mixed_rows(const Pattern &P, const MixedLU &m)
{
    if (threadIdx.x == 0) {
        const uint32_t n = P.n;
        for (uint32_t r = 0; r <= n; r++) m.rptr[r] = 0;
        for (uint32_t p = 0; p < P.nnz; p++) m.rptr[P.row_idx[p] + 1]++;
        for (uint32_t r = 0; r < n; r++) m.rptr[r + 1] += m.rptr[r];
        // fill: a running cursor per row kept in the row pointer of the next
        for (uint32_t c = 0; c < n; c++)
            for (uint32_t p = P.col_ptr[c]; p < P.col_ptr[c + 1]; p++) {
                const uint32_t r = P.row_idx[p];
                uint32_t k = m.rptr[r];
                m.rent[2 * k] = p;
                m.rent[2 * k + 1] = c;
                m.rptr[r] = k + 1;
            }
        for (uint32_t r = n; r > 0; r--) m.rptr[r] = m.rptr[r - 1];
        m.rptr[0] = 0;
    }
    __syncthreads();
}

// The wave's largest value, NaN counted as infinite so every lane agrees.
__device__ inline double wave_max_d(double v)
{
    if (!(v == v)) v = __builtin_huge_val();
    for (int s = 32; s >= 1; s >>= 1) {
        const double o = __shfl_xor(v, s, 64);
        v = o > v ? o : v;
    }
    return v;
}

// The power of two that brings v's magnitude into [0.5, 1): exact to
// multiply by in FP64; 1 for 0 and for non-finite v.
__device__ inline double pow2_scale(double v)
{
    if (!(v > 0.0) || !__builtin_isfinite(v)) return 1.0;
    int e = 0;
    (void)__builtin_frexp(v, &e);
    return __builtin_ldexp(1.0, -e);
}

// Factors vals (the lane's FP64 matrix, CSC) in FP32 with partial
// pivoting, after equilibrating it by powers of two (rows to a largest
// magnitude in [0.5, 1), then columns), exactly, in FP64: MNA rows mix
// conductances, source incidences and gmin-sized entries over many
// decades, which rounding to FP32 unscaled would make far worse
// conditioned than the FP64 matrix. P (R A C) = L U in m.F, the row at
// position k coming from row perm_k (kept by thread k); thread k also
// keeps row k's scale rs and column k's scale cs; `scratch` (n doubles of
// LDS) passes the row scales to the columns. False when a pivot is zero or
// any value is not finite in FP32.
__device__ bool mixed_factor(const Pattern &P, const MixedLU &m, const double *vals, uint32_t &perm_k, double &rs,
                             double &cs, double *scratch, double &anorm)
{
    const uint32_t n = P.n, tid = threadIdx.x;
    for (uint32_t i = tid; i < n * n; i += blockDim.x) m.F[i] = 0.0f;
    rs = 1.0;
    cs = 1.0;
    double rsum = 0.0;
    if (tid < n) {
        double rmax = 0.0;
        for (uint32_t e = m.rptr[tid]; e < m.rptr[tid + 1]; e++) {
            const double a = __builtin_fabs(vals[m.rent[2 * e]]);
            rmax = a > rmax ? a : rmax;
            rsum += a;
        }
        rs = pow2_scale(rmax);
        scratch[tid] = rs;
    }
    anorm = wave_max_d(rsum);
    __syncthreads();
    if (tid < n) {
        double cmax = 0.0;
        for (uint32_t p = P.col_ptr[tid]; p < P.col_ptr[tid + 1]; p++) {
            const double a = __builtin_fabs(vals[p]) * scratch[P.row_idx[p]];
            cmax = a > cmax ? a : cmax;
        }
        cs = pow2_scale(cmax);
    }
    bool bad = false;
    if (tid < n)
        for (uint32_t p = P.col_ptr[tid]; p < P.col_ptr[tid + 1]; p++) {
            const float v = (float)(vals[p] * scratch[P.row_idx[p]] * cs);
            bad = bad || !__builtin_isfinite(v);
            m.F[tid * n + P.row_idx[p]] = v;
        }
    perm_k = tid;
    if (wave_ballot(bad)) return false;
    __syncthreads();
    for (uint32_t k = 0; k < n; k++) {
        // the largest magnitude of column k at or below the diagonal, the
        // lowest row among equals
        // (a NaN counts as infinite, so every lane reaches the same choice)
        float best = (tid >= k && tid < n) ? __builtin_fabsf(m.F[k * n + tid]) : -1.0f;
        if (!(best == best)) best = __builtin_huge_valf();
        uint32_t at = tid;
        for (int s = 32; s >= 1; s >>= 1) {
            const float ob = __shfl_xor(best, s, 64);
            const uint32_t oa = __shfl_xor(at, s, 64);
            if (ob > best || (ob == best && oa < at)) {
                best = ob;
                at = oa;
            }
        }
        if (!(best > 0.0f) || !__builtin_isfinite(best)) return false;
        if (at != k) {
            if (tid < n) {
                const float a = m.F[tid * n + k], b = m.F[tid * n + at];
                m.F[tid * n + k] = b;
                m.F[tid * n + at] = a;
            }
            const uint32_t pk = __shfl(perm_k, k, 64), pa = __shfl(perm_k, at, 64);
            if (tid == k) perm_k = pa;
            if (tid == at) perm_k = pk;
        }
        __syncthreads();
        const float piv = m.F[k * n + k];
        if (tid > k && tid < n) m.F[k * n + tid] /= piv;
        __syncthreads();
        if (tid > k && tid < n) {
            const float u = m.F[tid * n + k];
            if (u != 0.0f)
                for (uint32_t r = k + 1; r < n; r++) m.F[tid * n + r] -= m.F[k * n + r] * u;
        }
        __syncthreads();
    }
    return true;
}

// y = U^-1 L^-1 (P b) in FP32; thread k passes in b's row k and gets back
// component k of the solution.
__device__ float mixed_substitute(const MixedLU &m, uint32_t n, uint32_t perm_k, double b_row)
{
    const uint32_t tid = threadIdx.x;
    float y = (float)__shfl(b_row, (int)perm_k, 64);
    for (uint32_t k = 0; k + 1 < n; k++) {
        const float yk = __shfl(y, (int)k, 64);
        if (tid > k && tid < n) y -= m.F[k * n + tid] * yk;
    }
    for (uint32_t k = n; k-- > 0;) {
        if (tid == k) y /= m.F[k * n + k];
        const float yk = __shfl(y, (int)k, 64);
        if (tid < k) y -= m.F[k * n + tid] * yk;
    }
    return y;
}

// The tick's linear step A dx = -rhs (vals: A in FP64, CSC): the FP32
// factor, then refinement with FP64 residuals r = -rhs - A dx until the
// residual meets the FP64 solve's own backward-error bar,
// max|r| <= max|dx| * ||A||_inf * sqrt(n) * 2^-53 (LAPACK dsgesv's
// criterion with BWDMAX = 1); dx lands in dx_out (s.dx). `solves` counts
// the FP32 solves this tick took. False when the factor fails, a value is
// not finite, or max_refine solves do not meet the bar; dx_out then holds
// scratch and the caller solves in FP64.
__device__ bool mixed_step(const Pattern &P, const MixedLU &m, const double *vals, const double *rhs, double *dx_out,
                           uint32_t max_refine, uint32_t &solves)
{
    const uint32_t n = P.n, tid = threadIdx.x;
    solves = 0;
    uint32_t perm_k;
    double rs, cs, anorm;
    if (!mixed_factor(P, m, vals, perm_k, rs, cs, dx_out, anorm)) return false;
    const double bar = anorm * __builtin_sqrt((double)n) * 0x1p-53;
    const double b = tid < n ? -rhs[tid] : 0.0;
    double r = b, dx = 0.0;
    bool ok = false;
    for (uint32_t it = 0; it < max_refine; it++) {
        // (R A C) z = R r, the correction C z
        const float y = mixed_substitute(m, n, perm_k, r * rs);
        solves++;
        dx += tid < n ? (double)y * cs : 0.0;
        // r = b - A dx in FP64, row by row
        __syncthreads();
        if (tid < n) dx_out[tid] = dx;
        __syncthreads();
        if (tid < n) {
            double acc = b;
            for (uint32_t e = m.rptr[tid]; e < m.rptr[tid + 1]; e++)
                acc -= vals[m.rent[2 * e]] * dx_out[m.rent[2 * e + 1]];
            r = acc;
        }
        const double rmax = wave_max_d(__builtin_fabs(r)), dmax = wave_max_d(__builtin_fabs(dx));
        if (!__builtin_isfinite(rmax) || !__builtin_isfinite(dmax)) break;
        if (rmax <= dmax * bar) {
            ok = true;
            break;
        }
    }
    __syncthreads();
    return ok;
}

#endif

}  // namespace esp
