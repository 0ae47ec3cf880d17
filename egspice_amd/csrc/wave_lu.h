// Wave-parallel replay of a lane's direct solver, for patterns of at most
// 64 unknowns: the lane's 64-wide wavefront runs the refactor and the two
// triangular solves with thread r owning pivot row r.
//
// The serial bodies in lane_solver.h (the reference's SparseLu replay,
// sparse_lu.zig:1066-1159, and luSolve) walk a column's U entries in the
// tape's topological order and, for each, subtract its L column from the
// scattered work vector. Here the work vector lives in registers (one entry
// per thread), each U value is broadcast with a lane read, and every L row
// of it updates its own entry at once. Each row still receives the same
// operations in the same order, and the mask of the row's structural
// entries keeps structural zeros out of it (no `w - 0*u`, whose sign of
// zero could differ), so the factor and the solution are bitwise the
// serial ones. Failure checks that the serial code makes column by column
// are deferred to one vote at the end: a failed replay's partial factor is
// never read (the caller re-pivots), so only the verdict has to agree.
//
// Values are kept dense in pivot coordinates, F[k*n + r] = L(r,k) for
// r > k, U(r,k) for r < k and the pivot for r == k; a thread reads and
// writes only its own row of F except for the pivots, so no barrier is
// needed inside the replay. The full factor (DFS reach, threshold pivot
// search) stays serial on thread 0 in lane_solver.h; after it the dense
// arrays are rebuilt from its tape.
#pragma once

#include "wave.h"

namespace esp {

constexpr uint16_t AMAP_NONE = 0xffffu;

struct DenseLU {
    double *F;        // n*n, column k at F + k*n, rows in pivot order
    uint64_t *lmask;  // n: bit k of lmask[r] when L(r,k) is structural
    uint64_t *umask;  // n: bit k of umask[r] when U(r,k) is structural
    uint16_t *amap;   // n*n: pattern slot of A at (pivot row r, step k), or AMAP_NONE
    uint32_t *porig;  // n: original row of pivot row t
    // the U column lists (the solver's ui, up[n] entries in its order) in
    // LDS when the solver keeps its own lists in HBM (WaveLayout.lu_global),
    // else null and the replay reads the solver's LDS lists
    uint32_t *ucols;
    const uint32_t *pred;  // the template's predicted first factor (PredTape, global memory), or null
    uint32_t *pred_hit;    // the lane's record of its first factor: 1 predicted, 2 serial (global), or null
    bool staged;           // pred_stage already copied the tape into the lane's arrays
};

// A first factor's symbolic tape, as the template's analysis writes it to
// global memory for every lane to try (dense_first_factor): header words
// then the arrays at the offsets below, all for n <= 64.
enum { PRED_OK = 0, PRED_NL = 1, PRED_NU = 2, PRED_HEAD = 8 };
struct PredTape {
    uint32_t pinv, lp, up, li, ui, prow, lmask, umask, amap, porig, words;
};
ESP_HD PredTape pred_tape(uint32_t n, uint32_t nnz)
{
    PredTape t;
    t.pinv = PRED_HEAD;
    t.lp = t.pinv + n;
    t.up = t.lp + n + 1;
    t.li = t.up + n + 1;
    t.ui = t.li + n * n;
    t.prow = t.ui + n * n;
    t.lmask = t.prow + nnz;
    t.umask = t.lmask + 2 * n;
    t.amap = t.umask + 2 * n;
    t.porig = t.amap + n * n;
    t.words = t.porig + n;
    return t;
}

ESP_HD DenseLU dense_lu(double *block, const WaveLayout &L, uint32_t n, const uint32_t *pred = nullptr,
                        uint32_t *pred_hit = nullptr)
{
    DenseLU d;
    d.pred = pred;
    d.pred_hit = pred_hit;
    d.staged = false;
    double *f = block + L.dense;
    d.F = f;
    f += (size_t)n * n;
    d.lmask = reinterpret_cast<uint64_t *>(f);
    f += n;
    d.umask = reinterpret_cast<uint64_t *>(f);
    f += n;
    d.amap = reinterpret_cast<uint16_t *>(f);
    f += ((size_t)n * n + 3) / 4;
    d.porig = reinterpret_cast<uint32_t *>(f);
    f += ((size_t)n + 1) / 2;
    d.ucols = L.lu_global ? reinterpret_cast<uint32_t *>(f) : nullptr;
    return d;
}

// amap entry i as a pattern slot, or ESP_NONE (slots fit 16 bits: nnz <= n*n
// <= 4096 for the dense replay's n <= 64).
__device__ inline uint32_t amap_slot(const DenseLU &D, uint32_t i)
{
    const uint32_t a = D.amap[i];
    return a == AMAP_NONE ? ESP_NONE : a;
}

// The U column lists the replay reads: the LDS copy, or the solver's own.
__device__ inline const uint32_t *dense_ucols(const DenseLU &D, const LaneLU &s)
{
    return D.ucols != nullptr ? D.ucols : s.ui.p;
}

// ----------------------------------------------------------- wave helpers

// Lane `src` (wave-uniform) of a per-thread double.
__device__ inline double lane_read(double v, uint32_t src)
{
    const uint64_t b = bits(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, (int)src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), (int)src);
    return from_bits((uint64_t)hi << 32 | lo);
}

__device__ inline uint64_t wave_ballot(bool p) { return __ballot(p ? 1 : 0); }

// ------------------------------------------------------------- the tape

// Dense arrays from the serial full factor's tape (sparse lx/li/lp,
// ux/ui/up, udiag, prow, pinv). All threads; ends with a barrier.
__device__ inline void dense_build(const Pattern &P, const LaneLU &s, const DenseLU &D)
{
    const uint32_t n = P.n, tid = threadIdx.x, nt = blockDim.x;
    for (uint32_t i = tid; i < n; i += nt) {
        D.lmask[i] = 0;
        D.umask[i] = 0;
        D.porig[s.pinv[i]] = i;
    }
    for (uint32_t i = tid; i < n * n; i += nt) D.amap[i] = AMAP_NONE;
    __syncthreads();
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = P.q[k];
        for (uint32_t p = s.lp[k] + tid; p < s.lp[k + 1]; p += nt) {
            const uint32_t r = s.li[p];
            D.F[k * n + r] = s.lx[p];
            atomicOr(reinterpret_cast<unsigned long long *>(&D.lmask[r]), 1ull << k);
        }
        for (uint32_t p = s.up[k] + tid; p < s.up[k + 1]; p += nt) {
            const uint32_t r = s.ui[p];
            D.F[k * n + r] = s.ux[p];
            atomicOr(reinterpret_cast<unsigned long long *>(&D.umask[r]), 1ull << k);
        }
        if (tid == 0) D.F[k * n + k] = s.udiag[k];
        for (uint32_t p = P.col_ptr[c] + tid; p < P.col_ptr[c + 1]; p += nt) D.amap[k * n + s.prow[p]] = (uint16_t)p;
    }
    if (D.ucols != nullptr)
        for (uint32_t p = tid; p < s.up[n]; p += nt) D.ucols[p] = s.ui[p];
    __syncthreads();
}

// ------------------------------------------------------------ the replay

// Lane `src` (wave-uniform) of a per-thread u32.
__device__ inline uint32_t lane_read_u32(uint32_t v, uint32_t src)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)src);
}

enum { REPLAY_HOLDS = 0, REPLAY_PROBE = 1, REPLAY_FAILED = 2 };

// Loads of CHUNK consecutive U entries (or solve columns) are issued back
// to back, unconditionally and in bounds, before their dependent updates,
// so the LDS round trip is paid once per chunk.
constexpr uint32_t CHUNK = 8;

// Replay.column on the lane's own pivot sequence (lane_solver.h
// lu_refactor), then replay_holds' multiplier bound: REPLAY_FAILED when
// the replay fails, REPLAY_HOLDS when every multiplier is within
// 1/pivot_tol (max_multiplier's columns), REPLAY_PROBE when the probe must
// decide. All threads; uniform result.
//
// A U row is final once consumed (every later entry of the column only
// updates rows below it in pivot order), so each thread keeps its row's
// value in `w` and stores it, with the L multipliers and the pivot, once
// at the end of the column. The next column's index list, scattered value
// and amap entry are loaded while this column runs.
template <class A>
__device__ int dense_refactor(const Pattern &P, const LaneLU &s, const DenseLU &D, const A &vals)
{
    const uint32_t *const ucol = dense_ucols(D, s);
    const uint32_t n = P.n, r = threadIdx.x, nt = blockDim.x;
    const bool own = r < n;
    const uint32_t rr = own ? r : n - 1;  // in-bounds row for unconditional loads
    const uint32_t nvoid = s.cnt[C_NVOID];
    bool live_void = false;
    for (uint32_t t = r; t < nvoid; t += nt) live_void = live_void || vals[s.void_slots[t]] != 0;
    if (wave_ballot(live_void)) return REPLAY_FAILED;
    const uint64_t lm = own ? D.lmask[r] : 0, um = own ? D.umask[r] : 0;
    const uint64_t scaled = wave_ballot(own && s.scaled_pivot[r] != 0);
    const uint64_t voided = wave_ballot(own && s.void_col[r] != 0);
    // thread k holds column k's U range
    const uint32_t ub = own ? s.up[r] : 0, ue = own ? s.up[r + 1] : 0;
    uint64_t nonfinite = 0;
    bool growth_fail = false, pivot_fail = false, wide_all = false, wide_kept = false, nan_l = false;
    const uint32_t a0 = own ? amap_slot(D, r) : ESP_NONE;
    uint32_t a_next = own && n > 1 ? amap_slot(D, n + r) : ESP_NONE;
    double wv = a0 != ESP_NONE ? vals[a0] : 0.0;
    uint32_t u0 = lane_read_u32(ub, 0), u1 = lane_read_u32(ue, 0);
    uint32_t uidx = r < u1 - u0 ? ucol[u0 + r] : 0;
    for (uint32_t k = 0; k < n; k++) {
        double w = wv;
        const uint32_t cnt = u1 - u0, col_idx = uidx;
        if (k + 1 < n) {
            wv = a_next != ESP_NONE ? vals[a_next] : 0.0;
            a_next = own && k + 2 < n ? amap_slot(D, (k + 2) * n + r) : ESP_NONE;
            u0 = lane_read_u32(ub, k + 1);
            u1 = lane_read_u32(ue, k + 1);
            uidx = r < u1 - u0 ? ucol[u0 + r] : 0;
        }
        for (uint32_t j0 = 0; j0 < cnt; j0 += CHUNK) {
            uint32_t iv[CHUNK];
            double fv[CHUNK];
#pragma unroll
            for (uint32_t c = 0; c < CHUNK; c++) {
                const uint32_t j = j0 + c < cnt ? j0 + c : cnt - 1;
                iv[c] = lane_read_u32(col_idx, j);
                fv[c] = D.F[iv[c] * n + rr];
            }
#pragma unroll
            for (uint32_t c = 0; c < CHUNK; c++) {
                if (j0 + c < cnt) {
                    const double uki = lane_read(w, iv[c]);
                    const double upd = w - fv[c] * uki;
                    w = (lm >> iv[c] & 1) ? upd : w;
                }
            }
        }
        const bool is_u = um >> k & 1, is_l = lm >> k & 1;
        if (is_u) {
            D.F[k * n + r] = w;
            nonfinite |= bits(w - w);
        }
        if (voided >> k & 1) {
            if (r == k) D.F[k * n + k] = 1;
            if (is_l) D.F[k * n + r] = 0;
            continue;
        }
        const double d = lane_read(w, k);
        if (r == k) D.F[k * n + k] = d;
        pivot_fail = pivot_fail || d == 0 || !isfinite_(d);
        if (is_l) {
            const double l = w / d;
            D.F[k * n + r] = l;
            const double al = fabs_(l);
            const bool wide = !(al <= 1 / PIVOT_TOL);
            nan_l = nan_l || al != al;
            wide_all = wide_all || wide;
            if (!(scaled >> k & 1)) {
                // growth: |d| against every unscaled step's column entry
                growth_fail = growth_fail || !(fabs_(d) >= GROWTH_LIMIT * fabs_(w));
                wide_kept = wide_kept || wide;
            } else {
                nonfinite |= bits(l - l);
            }
        }
    }
    if (pivot_fail || wave_ballot(growth_fail || nonfinite != 0) != 0) return REPLAY_FAILED;
    // max_multiplier <= 1/pivot_tol (lane_solver.h): every multiplier of L
    // within the bound, or, when some pivot was taken by the scaled test and
    // no multiplier is NaN, every one of the other columns
    const bool wide = (scaled == 0 || wave_ballot(nan_l)) ? wave_ballot(wide_all) != 0 : wave_ballot(wide_kept) != 0;
    return wide ? REPLAY_PROBE : REPLAY_HOLDS;
}

// lu_solve's two sweeps on the wavefront: thread r enters with y_r, the
// right-hand side at pivot position r, and leaves with the solution's
// entry there. Forward over L skips a zero y_k, backward over U a zero
// z_k, as the serial sweeps do. All threads; reads the factor other
// threads wrote, so it must follow a barrier.
__device__ inline double dense_sweeps(uint32_t n, const DenseLU &D, double y)
{
    const uint32_t r = threadIdx.x;
    const bool own = r < n;
    const uint32_t rr = own ? r : n - 1;
    const uint64_t lm = own ? D.lmask[r] : 0, um = own ? D.umask[r] : 0;
    for (uint32_t k0 = 0; k0 < n; k0 += CHUNK) {
        double fv[CHUNK];
#pragma unroll
        for (uint32_t c = 0; c < CHUNK; c++) {
            const uint32_t k = k0 + c < n ? k0 + c : n - 1;
            fv[c] = D.F[k * n + rr];
        }
#pragma unroll
        for (uint32_t c = 0; c < CHUNK; c++) {
            const uint32_t k = k0 + c;
            if (k < n) {
                const double yk = lane_read(y, k);
                const double upd = y - fv[c] * yk;
                y = (yk != 0 && (lm >> k & 1)) ? upd : y;
            }
        }
    }
    for (uint32_t k1 = n; k1 > 0;) {
        double fv[CHUNK], dv[CHUNK];
#pragma unroll
        for (uint32_t c = 0; c < CHUNK; c++) {
            const uint32_t k = k1 > c ? k1 - 1 - c : 0;
            fv[c] = D.F[k * n + rr];
            dv[c] = D.F[k * n + k];
        }
#pragma unroll
        for (uint32_t c = 0; c < CHUNK; c++) {
            if (k1 > c) {
                const uint32_t k = k1 - 1 - c;
                const double zk = lane_read(y, k) / dv[c];
                const double upd = y - fv[c] * zk;
                y = r == k ? zk : ((zk != 0 && (um >> k & 1)) ? upd : y);
            }
        }
        k1 = k1 > CHUNK ? k1 - CHUNK : 0;
    }
    return y;
}

// luSolve of -rhs into dx (lane_solver.h direct_solve_neg / lu_solve).
// All threads; reads the pivots other threads wrote, so it must follow a
// barrier; writes dx[q[r]] (needs a barrier before others read it).
template <class R, class X>
__device__ void dense_solve_neg(const Pattern &P, const DenseLU &D, const R &rhs, X &dx)
{
    const uint32_t n = P.n, r = threadIdx.x;
    const bool own = r < n;
    double y = own ? -1.0 * rhs[D.porig[r]] : 0.0;
    y = dense_sweeps(n, D, y);
    if (own) dx[P.q[r]] = y;
}

// lu_solve_t's sweeps (U^T then L^T) on the wavefront, thread r holding
// y_r at pivot position r. Each step's sum runs over the column's U (L)
// rows in the tape's order, one term per row, as the serial dot product
// subtracts them; the products are formed on their own rows in parallel.
__device__ inline double dense_sweeps_t(uint32_t n, const LaneLU &s, const DenseLU &D, double y)
{
    const uint32_t *const ucol = dense_ucols(D, s);
    const uint32_t r = threadIdx.x;
    const uint32_t rr = r < n ? r : n - 1;
    for (uint32_t k = 0; k < n; k++) {
        const double prod = D.F[k * n + rr] * y;
        double v = lane_read(y, k);
        for (uint32_t p = s.up[k]; p < s.up[k + 1]; p++) v = v - lane_read(prod, ucol[p]);
        const double yk = v / D.F[k * n + k];
        y = r == k ? yk : y;
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        const double prod = D.F[k * n + rr] * y;
        double v = lane_read(y, k);
        for (uint32_t p = s.lp[k]; p < s.lp[k + 1]; p++) v = v - lane_read(prod, s.li[p]);
        y = r == k ? v : y;
    }
    return y;
}

// ------------------------------------------------------- replay's probe

// The largest bit pattern over the wavefront (norm_inf_lv's bitwise max;
// callers pass 0 from idle threads).
__device__ inline uint64_t wave_bitmax(uint64_t b)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const uint64_t o = __shfl_xor(b, m, 64);
        b = o > b ? o : b;
    }
    return b;
}

// probe_rhs's entry i (lane_solver.h).
__device__ inline double probe_entry(uint32_t i, uint32_t n)
{
    const double den = (double)((n > 2 ? n : 2) - 1);
    const double mag = 1 + (double)i / den;
    return i % 2 == 0 ? mag : -mag;
}

// dense_probe_holds on the wavefront: thread i forms row i's sums and
// residual in the serial order (columns ascending), thread j column j's;
// the norms are order-free bit maxima; the two solves are dense_sweeps and
// dense_sweeps_t. `px` is n doubles of LDS for the probe's solution. All
// threads; uniform result.
template <class A>
__device__ bool dense_probe_holds_wave(const Pattern &P, const LaneLU &s, const DenseLU &D, const A &vals, double *px)
{
    const uint32_t n = P.n, r = threadIdx.x;
    const bool own = r < n;
    double row = 0, col = 0;
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = P.col_ptr[j]; p < P.col_ptr[j + 1]; p++)
            if (P.row_idx[p] == r) row += fabs_(vals[p]);
    if (own)
        for (uint32_t p = P.col_ptr[r]; p < P.col_ptr[r + 1]; p++) col += fabs_(vals[p]);
    // `if (col > norm_1) norm_1 = col` from 0: the largest non-NaN column sum
    const double norm_1 = from_bits(wave_bitmax(own && col == col ? bits(col) : 0));
    const double norm_inf_a = from_bits(wave_bitmax(own ? bits(fabs_(row)) : 0));
    const double nf = (double)n;
    for (int transpose = 0; transpose < 2; transpose++) {
        if (!transpose) {
            const double y = dense_sweeps(n, D, own ? probe_entry(D.porig[r], n) : 0.0);
            if (own) px[P.q[r]] = y;
        } else {
            const double y = dense_sweeps_t(n, s, D, own ? probe_entry(P.q[r], n) : 0.0);
            if (own) px[D.porig[r]] = y;
        }
        __syncthreads();
        double res = own ? probe_entry(r, n) : 0.0;
        if (!transpose) {
            for (uint32_t j = 0; j < n; j++) {
                const double xj = px[j];
                for (uint32_t p = P.col_ptr[j]; p < P.col_ptr[j + 1]; p++)
                    if (P.row_idx[p] == r) res -= vals[p] * xj;
            }
        } else if (own) {
            for (uint32_t p = P.col_ptr[r]; p < P.col_ptr[r + 1]; p++) res -= vals[p] * px[P.row_idx[p]];
        }
        const double nr = from_bits(wave_bitmax(own ? bits(fabs_(res)) : 0));
        const double nx = from_bits(wave_bitmax(own ? bits(fabs_(px[r])) : 0));
        const double a_norm = transpose ? norm_1 : norm_inf_a;
        const double ratio = nr / (a_norm * nx * nf * EPS);
        __syncthreads();
        if (!(ratio <= 30.0 / PIVOT_TOL)) return false;
    }
    return true;
}

// ------------------------------------------- the predicted first factor

// The predicted tape's symbolic arrays into the lane's solver (pinv, lp,
// up, li, ui, prow) and dense arrays (porig, masks, amap). It reads no
// matrix value, so a kernel may run it in its prologue, its loads
// overlapping the template's; the index lists are copied to the solver's
// capacity (only a tape within it is ever used). All threads; no barrier.
__device__ inline void pred_stage(const Pattern &P, LaneLU &s, const DenseLU &D)
{
    const uint32_t *g = D.pred;
    if (g == nullptr) return;
    const uint32_t n = P.n, nnz = P.nnz, r = threadIdx.x, nt = blockDim.x;
    const PredTape T = pred_tape(n, nnz);
    const uint32_t lists = n * n < P.cap ? n * n : (uint32_t)P.cap;
    if (r < n) {
        s.pinv[r] = g[T.pinv + r];
        D.porig[r] = g[T.porig + r];
        D.lmask[r] = (uint64_t)g[T.lmask + 2 * r + 1] << 32 | g[T.lmask + 2 * r];
        D.umask[r] = (uint64_t)g[T.umask + 2 * r + 1] << 32 | g[T.umask + 2 * r];
    }
    for (uint32_t i = r; i <= n; i += nt) {
        s.lp[i] = g[T.lp + i];
        s.up[i] = g[T.up + i];
    }
    for (uint32_t i = r; i < lists; i += nt) {
        s.li[i] = g[T.li + i];
        s.ui[i] = g[T.ui + i];
    }
    if (D.ucols != nullptr) {
        const uint32_t nu_ = g[T.up + n];
        for (uint32_t i = r; i < nu_ && i < lists; i += nt) D.ucols[i] = g[T.ui + i];
    }
    for (uint32_t i = r; i < nnz; i += nt) s.prow[i] = g[T.prow + i];
    for (uint32_t i = r; i < n * n; i += nt) D.amap[i] = g[T.amap + i] == ESP_NONE ? AMAP_NONE : (uint16_t)g[T.amap + i];
}

// The lane's first full factor (lane_solver.h lu_factor on a fresh solver:
// factor_pass with keep_prev off) computed on the wavefront from the
// template's predicted tape. factor_pass's DFS, reach and operation order
// depend only on the pattern and on the pivot chosen at each earlier
// step, so as long as every step's choice equals the prediction, the
// prediction's tape is the one factor_pass would build and this replay's
// values are its values, bit for bit (dense_refactor's argument). At each
// step the candidates (the pivot row and the column's L rows) are judged
// by factor_pass's rule: the largest magnitude among them (first in
// topological order on a tie), unless the wanted row want[c] passes the
// threshold test, plain or row-scaled (which marks the step scaled); the
// wanted rows follow factor_pass's swaps. Returns true with the solver in
// exactly the state lu_factor leaves; false (nothing the serial factor
// reads changed) when the prediction is absent or any step would choose
// otherwise, and on a tie, a zero or non-finite candidate, a void column,
// or a fill over the cap, all of which the serial factor then decides.
// All threads; uniform result.
template <class A>
__device__ bool dense_first_factor(const Pattern &P, LaneLU &s, const DenseLU &D, const A &vals)
{
    const uint32_t *const ucol = dense_ucols(D, s);
    const uint32_t n = P.n, nnz = P.nnz, r = threadIdx.x, nt = blockDim.x;
    const uint32_t *g = D.pred;
    if (g == nullptr || g[PRED_OK] != 1) return false;
    const uint32_t nl = g[PRED_NL], nu = g[PRED_NU];
    uint64_t cap = P.cap;
    if (P.diag_fill > 0) {
        const double c = 3.0 * (double)P.diag_fill;
        const uint64_t cc = (uint64_t)(c < 1e18 ? c : 1e18);
        if (cc < cap) cap = cc;
    }
    if ((uint64_t)nl + nu > cap) return false;
    const bool own = r < n;
    const uint32_t rr = own ? r : n - 1;
    // factor_pass's opening checks: every value finite, then the row
    // scales, each row's largest magnitude (a bit-pattern maximum, exact
    // for finite non-negative values) gathered column by column
    uint64_t mb = 0;
    for (uint32_t p = r; p < nnz; p += nt) {
        const uint64_t b = bits(fabs_(vals[p]));
        mb = b > mb ? b : mb;
    }
    if (!(from_bits(wave_bitmax(mb)) <= 1.7976931348623157e308)) return false;
    unsigned long long *const rbits = reinterpret_cast<unsigned long long *>(&s.rscale[0]);
    if (own) rbits[r] = 0;
    __syncthreads();
    if (own)
        for (uint32_t p = P.col_ptr[r]; p < P.col_ptr[r + 1]; p++) atomicMax(&rbits[P.row_idx[p]], bits(fabs_(vals[p])));
    __syncthreads();
    // the predicted tape into the lane's solver and dense arrays
    if (own) {
        const double rmax = from_bits(rbits[r]);
        s.rscale[r] = (rmax > 0 && isfinite_(rmax)) ? 1 / rmax : 1;
        s.void_col[r] = 0;
        s.scaled_pivot[r] = 0;
    }
    if (!D.staged) pred_stage(P, s, D);
    __syncthreads();
    // the original row of every pivot row, and the pivot row of every
    // original row, one per thread
    const uint32_t orig = own ? D.porig[r] : 0, pos_of = own ? s.pinv[r] : ESP_NONE;
    const double rs_orig = own ? s.rscale[orig] : 1.0;  // rscale of pivot row r's original row
    uint32_t want = r, want_of = r;                      // factor_pass's want / want_of, thread i holding entry i
    const uint64_t lm = own ? D.lmask[r] : 0, um = own ? D.umask[r] : 0;
    const uint32_t ub = own ? s.up[r] : 0, ue = own ? s.up[r + 1] : 0;
    bool bad = false;
    uint64_t scaled_steps = 0;
    const uint32_t a0 = own ? amap_slot(D, r) : ESP_NONE;
    uint32_t a_next = own && n > 1 ? amap_slot(D, n + r) : ESP_NONE;
    double wv = a0 != ESP_NONE ? vals[a0] : 0.0;
    uint32_t u0 = lane_read_u32(ub, 0), u1 = lane_read_u32(ue, 0);
    uint32_t uidx = r < u1 - u0 ? ucol[u0 + r] : 0;
    for (uint32_t k = 0; k < n && !bad; k++) {
        double w = wv;
        const uint32_t cnt = u1 - u0, col_idx = uidx;
        if (k + 1 < n) {
            wv = a_next != ESP_NONE ? vals[a_next] : 0.0;
            a_next = own && k + 2 < n ? amap_slot(D, (k + 2) * n + r) : ESP_NONE;
            u0 = lane_read_u32(ub, k + 1);
            u1 = lane_read_u32(ue, k + 1);
            uidx = r < u1 - u0 ? ucol[u0 + r] : 0;
        }
        for (uint32_t j0 = 0; j0 < cnt; j0 += CHUNK) {
            uint32_t iv[CHUNK];
            double fv[CHUNK];
#pragma unroll
            for (uint32_t c = 0; c < CHUNK; c++) {
                const uint32_t j = j0 + c < cnt ? j0 + c : cnt - 1;
                iv[c] = lane_read_u32(col_idx, j);
                fv[c] = D.F[iv[c] * n + rr];
            }
#pragma unroll
            for (uint32_t c = 0; c < CHUNK; c++) {
                if (j0 + c < cnt) {
                    const double uki = lane_read(w, iv[c]);
                    const double upd = w - fv[c] * uki;
                    w = (lm >> iv[c] & 1) ? upd : w;
                }
            }
        }
        const bool is_u = um >> k & 1, is_l = lm >> k & 1, cand = own && (r == k || is_l);
        if (is_u) D.F[k * n + r] = w;
        // every reach row's value finite (factor_pass's nonfinite), then
        // the choice among the candidates
        const double a = fabs_(w);
        const bool nonfinite = (is_u || cand) && !(a < inf_());
        const uint64_t abits = cand ? bits(a) : 0;
        const uint64_t amax_b = wave_bitmax(abits);
        const uint64_t at_max = wave_ballot(cand && abits == amax_b);
        const double amax = from_bits(amax_b);
        const uint32_t c = P.q[k];
        const uint32_t pr = lane_read_u32(want, c);            // want[c], an original row
        const uint32_t prk = lane_read_u32(pos_of, pr);        // its predicted pivot row
        const uint64_t cands = wave_ballot(cand);
        uint32_t chosen = (uint32_t)__builtin_ctzll(at_max | (1ull << 63));
        bool scaled_k = false, by_want = false;
        if (prk >= k && prk < n) {
            const double dmag = (cands >> prk & 1) ? fabs_(lane_read(w, prk)) : 0.0;
            const double rs_pr = lane_read(rs_orig, prk);
            if (dmag > 0 && dmag >= PIVOT_TOL * amax) {
                by_want = true;
            } else if (dmag > 0) {
                const double smax = from_bits(wave_bitmax(cand ? bits(a * rs_orig) : 0));
                if (dmag * rs_pr >= PIVOT_TOL * smax) {
                    by_want = true;
                    scaled_k = true;
                }
            }
        }
        if (by_want) chosen = prk;
        // a tie for the largest is settled by topological order, which the
        // dense rows do not keep: the serial factor decides it
        if (wave_ballot(nonfinite) != 0 || !(amax > 0) || chosen != k ||
            (!by_want && __builtin_popcountll(at_max) != 1)) {
            bad = true;
            break;
        }
        if (scaled_k) scaled_steps |= 1ull << k;
        // factor_pass's swap of the wanted rows when the pivot is not want[c]
        const uint32_t piv = lane_read_u32(orig, k);
        if (piv != pr) {
            const uint32_t c2 = lane_read_u32(want_of, piv);
            if (r == c2) want = pr;
            if (r == pr) want_of = c2;
        }
        const double d = lane_read(w, k);
        if (r == k) D.F[k * n + k] = d;
        if (is_l) {
            const double l = w / d;
            D.F[k * n + r] = l;
            bad = bad || !(fabs_(l) < inf_());
        }
        bad = wave_ballot(bad) != 0;
    }
    if (bad) return false;
    // the counters and flags lu_factor leaves; the factor's values live in
    // F (the serial path's sparse lx/ux/udiag only feed dense_build, and a
    // later full factor rewrites them before reading)
    if (own) {
        s.scaled_pivot[r] = (scaled_steps >> r & 1) ? 1 : 0;
        if (P.diag_fill > 0) s.prev[r] = r;
    }
    for (uint32_t p = r; p < nnz; p += nt) s.vcopy[p] = vals[p];
    if (r == 0) {
        s.cnt[C_NL] = nl;
        s.cnt[C_NU] = nu;
        s.cnt[C_NVOID] = 0;
        s.cnt[C_BASE] = nl + nu;
        s.state[ST_LU_FACTORED] = 1;
        s.state[ST_FACTORED] = 1;
    }
    __syncthreads();
    return true;
}

// --------------------------------------------------- direct.Solver.factor

// lane_solver.h direct_factor for one lane on its wavefront: 0 factored,
// 1 singular. `vals` is the lane's G plane in LDS; `flag` one LDS word for
// thread 0's verdicts. All threads; uniform result.
template <class A>
__device__ int wave_direct_factor(const Pattern &P, LaneLU &s, const DenseLU &D, const A &vals, uint32_t *flag)
{
    const uint32_t nnz = P.nnz, tid = threadIdx.x, nt = blockDim.x;
    __syncthreads();
    if (nnz != 0 && s.state[ST_FACTORED]) {
        bool differs = false;
        for (uint32_t p = tid; p < nnz; p += nt) differs = differs || !(vals[p] == s.vcopy[p]);
        if (!wave_ballot(differs)) return 0;
    }
    if (s.state[ST_TRI]) {
        // the Thomas engine, or its one-time demotion to the LU, serially
        if (tid == 0) *flag = (uint32_t)direct_factor(P, s, vals);
        __syncthreads();
        const uint32_t r = *flag;
        if (r == 0 && !s.state[ST_TRI]) dense_build(P, s, D);
        __syncthreads();
        return (int)r;
    }
    if (s.state[ST_FACTORED]) {
        const int replay = dense_refactor(P, s, D, vals);
        bool holds = replay == REPLAY_HOLDS;
        if (replay == REPLAY_PROBE) {
            __syncthreads();
            holds = dense_probe_holds_wave(P, s, D, vals, s.px.p);
        }
        if (holds) {
            for (uint32_t p = tid; p < nnz; p += nt) s.vcopy[p] = vals[p];
            __syncthreads();
            return 0;
        }
        __syncthreads();
        if (tid == 0) s.state[ST_FACTORED] = 0;
    }
    __syncthreads();
    if (!s.state[ST_FACTORED] && !s.state[ST_LU_FACTORED] && D.pred != nullptr) {
        const bool hit = dense_first_factor(P, s, D, vals);
        if (tid == 0 && D.pred_hit != nullptr && *D.pred_hit == 0) *D.pred_hit = hit ? 1u : 2u;
        if (hit) return 0;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t r = (uint32_t)lu_factor(P, s, vals);
        if (r == 0) s.state[ST_FACTORED] = 1;
        *flag = r;
    }
    __syncthreads();
    if (*flag) return 1;
    for (uint32_t p = tid; p < nnz; p += nt) s.vcopy[p] = vals[p];
    dense_build(P, s, D);
    return 0;
}

// update_norm (engine.h) across the wavefront: x_old = x, x += dx. The
// caller only compares the worst scaled step with 1 (finalizeStep's delta
// gate), so this returns 1 when some score reaches 1 (max >= 1 iff some
// element >= 1; the scores are non-negative, NaN scores count as 0) and 0
// otherwise, which the gate reads the same way as the max itself.
__device__ inline double wave_update_norm(uint32_t n, double *x, const double *dx, double *x_old, const uint8_t *cur_row,
                                          double reltol, double abstol, double vntol)
{
    bool big = false;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        double xo = x[i], d = dx[i], xn = xo + d;
        x_old[i] = xo;
        x[i] = xn;
        double tol = reltol * fmax_(fabs_(xn), fabs_(xo)) + (cur_row[i] ? abstol : vntol);
        bool finite = fabs_(xn) < inf_() && fabs_(d) < inf_();
        double q = fabs_(d) / tol;
        double score = !finite ? inf_() : (q == q ? q : 0);
        big = big || score >= 1.0;
    }
    return wave_ballot(big) ? 1.0 : 0.0;
}

}  // namespace esp
