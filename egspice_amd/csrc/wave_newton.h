// The lane's Newton on its wavefront: converger.newton (src/solver/
// converger.zig:155-319, finalizeStep :458-526) with the phases of wave.h
// and the dense replay of wave_lu.h, shared by the operating-point kernel
// and the transient (where it runs under the integrator's hook).
#pragma once

#include "wave_lu.h"
#include "wave_mixed.h"

namespace esp {

// integrator.Hook (src/analysis/tran/integrator.zig): the companion
// residual rhs += ag0 (q - q1) [- ag1 i_prev | - ag2 (q1 - q2)] on every
// row and the matrix a = G + ag0 C, added after the ground stamp.
struct WaveHook {
    int method;  // 0 BE, 1 trapezoidal, 2 gear-2
    double ag0, ag1, ag2;
    const double *q1, *q2, *i_prev;
    double *a;
};

enum { NEWTON_SINGULAR = -1, NEWTON_FAILED = 0, NEWTON_CONVERGED = 1 };

// The mixed-precision linear step's controls and tallies (wave_mixed.h):
// at most max_refine FP32 solves per tick; solves and the ticks handed back
// to the FP64 factor are counted per lane.
struct MixedRun {
    MixedLU m;
    uint32_t max_refine;
    uint32_t solves, fallbacks;
};

// Solves from s.x; a converged lane returns x_k (x = x_old) and the
// iteration count. TRAN: sources read their waveform at t (a transient's
// operating point and its steps); HOOK: the integrator's companion applies
// (h). Both are compile-time so the static solve carries none of it.
// MIXED: each tick's linear step runs in FP32 with FP64 refinement (mx).
// All threads; uniform result.
template <bool TRAN, bool HOOK, int VA = 0, bool MIXED = false>
__device__ __forceinline__ int wave_newton(const WaveTemplate &W, WaveLane &s, const DenseLU &D, const ReduceCache &rc, const Newton &o,
                           const WaveHook &h, bool init_fix, bool lim_active, double t, uint32_t &iters,
                           MixedRun *mx = nullptr)
{
    static_assert(TRAN || !HOOK, "the companion only runs in a transient");
    const uint32_t tid = threadIdx.x, nt = blockDim.x, n = W.P.n, nnz = W.P.nnz;
    uint32_t bits = init_fix ? G_INITFIX : 0;
    double limit = 1e9, scaled = 0;
    iters = o.max_iter;
    for (uint32_t i = tid; i < n; i += nt) s.x_old[i] = s.x[i];
    __syncthreads();
    for (uint32_t tick = 0; tick < o.max_iter; tick++) {
        for (uint32_t e = tid; e < W.n_inst; e += nt) wave_eval_inst<TRAN, VA>(W, s, e, lim_active, t);
        __syncthreads();
        wave_reduce(W, s, rc, tid, nt);
        if constexpr (HOOK)
            for (uint32_t e = tid; e < n + nnz; e += nt) wave_reduce_charge(W, s, e);
        __syncthreads();
        double *const vals = HOOK ? h.a : s.g;
        if constexpr (HOOK) {
            if (tid == 0) {
                s.g[W.diag[0]] += 1.0;
                s.rhs[0] += s.x[0];
            }
            __syncthreads();
            for (uint32_t i = tid; i < n; i += nt) {
                const double d = h.ag0 * (s.q[i] - h.q1[i]);
                const double base = s.rhs[i] + d;
                s.rhs[i] = h.method == 1 ? base - h.ag1 * h.i_prev[i]
                           : h.method == 2 ? base - h.ag2 * (h.q1[i] - h.q2[i])
                                           : base;
            }
            for (uint32_t p = tid; p <= nnz; p += nt) h.a[p] = s.g[p] + h.ag0 * (p < nnz ? s.cc[p] : 0.0);
            __syncthreads();
            if (tid == 0 && tick == 0) {
                LV<double> x{s.x, 1}, rhs{s.rhs, 1}, a{vals, 1};
                limit = plausible_limit(n, x, rhs, a, W.diag);
            }
        } else {
            if (tid == 0) wave_assemble(W, s, o, tick, limit);
        }
        bool mixed_done = false;
        if constexpr (MIXED) {
            uint32_t solves = 0;
            mixed_done = mixed_step(W.P, mx->m, vals, s.rhs, s.dx, mx->max_refine, solves);
            mx->solves += solves;
            if (!mixed_done) mx->fallbacks++;
        }
        const int fl = mixed_done ? 0 : wave_direct_factor(W.P, s.lu, D, vals, &s.flags[2]);
        if (mixed_done) {
            scaled = wave_update_norm(n, s.x, s.dx, s.x_old, W.cur_row, o.reltol, o.abstol, o.vntol);
        } else if (fl == 0) {
            if (s.lu.state[ST_TRI]) {
                if (tid == 0) {
                    LV<double> rhs{s.rhs, 1}, dx{s.dx, 1};
                    direct_solve_neg(W.P, s.lu, rhs, dx);
                }
            } else {
                dense_solve_neg(W.P, D, s.rhs, s.dx);
            }
            __syncthreads();
            scaled = wave_update_norm(n, s.x, s.dx, s.x_old, W.cur_row, o.reltol, o.abstol, o.vntol);
        }
        if (fl != 0) {
            iters = 0;
            return NEWTON_SINGULAR;
        }
        int lim_any = 0;
        for (uint32_t e = tid; e < W.n_inst; e += nt) lim_any |= wave_limit_inst<VA>(W, s, e, lim_active) ? 1 : 0;
        lim_any = __syncthreads_or(lim_any);
        lim_active = true;
        if (tid == 0) {
            bool ok = wave_gates(W, s, o, tick, lim_any != 0, limit, scaled, vals);
            uint32_t verdict = 0;
            if (ok && (bits & G_INITFIX)) bits &= ~G_INITFIX;
            else if (ok) verdict = 1;
            s.flags[1] = verdict;
        }
        __syncthreads();
        if (s.flags[1]) {
            for (uint32_t i = tid; i < n; i += nt) s.x[i] = s.x_old[i];
            iters = tick + 1;
            __syncthreads();
            return NEWTON_CONVERGED;
        }
        __syncthreads();
    }
    return NEWTON_FAILED;
}

}  // namespace esp
