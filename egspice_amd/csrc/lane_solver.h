// Per-lane sparse direct solver of the lane engine, one body for the host
// and the gfx950 kernels. Every array is a lane view (common.h LV): on the
// device lane l of N reads element i at base[i * N + l], so the lanes of a
// wavefront that take the same branch issue coalesced loads.
//
// A lane runs the reference CPU path's `direct.Solver.factor` on its own
// values each Newton tick (src/solver/direct.zig:310-383):
//   bypass when the values equal the last factored ones;
//   the Thomas engine while the pattern is tridiagonal (tridiag.zig:100-137),
//     demoted to the LU for good at its first failed pivot (direct.zig:359);
//   otherwise refactor on the lane's own pivot sequence (sparse_lu.zig
//     Replay.column :1066-1159) and keep it while `replayHolds` (direct.zig
//     :402-447: max multiplier within 1/pivot_tol, else a probe right-hand
//     side solved both ways within the LAWN 41 bar);
//   else a full left-looking Gilbert-Peierls factor (sparse_lu.zig:274-663):
//     DFS reach through L, triangular solve in reverse finish order,
//     threshold pivoting with the diagonal first, the row-scaled second
//     chance and the want hand-off, void unknowns, fill caps (3x the
//     diagonal-pivot fill, direct.zig:36) with the diagonal/previous-pivot
//     restart (sparse_lu.zig:274-311).
// Also here: the lead-tape bodies of the reference's GPU lanes
// (src/solver/kernels/lane.zig:56-157, `LaneTab`), replayed by
// `espice_lanes_tab_refactor_solve`.
#pragma once

#include "common.h"

namespace esp {

constexpr double PIVOT_TOL = 1e-3;
constexpr double GROWTH_LIMIT = 1e-12;
constexpr double EPS = 2.220446049250313e-16;

// Shared (per template) pattern data, read by every lane.
struct Pattern {
    uint32_t n, nnz;
    const uint32_t *col_ptr;  // n + 1
    const uint32_t *row_idx;  // nnz
    const uint32_t *q;        // n: BTF+AMD column order
    const uint32_t *tri_pos;  // 3n: a, b, c slots when tridiagonal, else null
    uint64_t diag_fill;       // max(order.diagFill, nnz)
    uint32_t cap;             // L+U capacity per lane
};

// Lane-major storage of every lane's solver (device or host pointers).
struct LaneLUBuf {
    uint8_t *state;  // 3 per lane: factored, tri engine, lu factored
    uint32_t *cnt;   // 4 per lane: nnz_l, nnz_u, n_void_slots, base_lu
    uint32_t *pinv, *lp, *up, *li, *ui, *prow, *void_slots;
    uint32_t *flag, *topo, *stack, *pstack, *want, *want_of, *prev;
    uint8_t *void_col, *scaled_pivot;
    double *lx, *ux, *udiag, *w, *y, *rscale, *vcopy, *tri, *probe;
};

// One lane's views into a LaneLUBuf.
struct LaneLU {
    LV<uint8_t> state;
    LV<uint32_t> cnt, pinv, lp, up, li, ui, prow, void_slots, flag, topo, stack, pstack, want, want_of, prev;
    LV<uint8_t> void_col, scaled_pivot;
    LV<double> lx, ux, udiag, w, y, rscale, vcopy, bp, mul, cv, px, pr, prows;
};

ESP_HD LaneLU lane_lu(const LaneLUBuf &b, size_t N, size_t l, uint32_t n)
{
    LaneLU s;
    s.state = lane_view(b.state, N, l);
    s.cnt = lane_view(b.cnt, N, l);
    s.pinv = lane_view(b.pinv, N, l);
    s.lp = lane_view(b.lp, N, l);
    s.up = lane_view(b.up, N, l);
    s.li = lane_view(b.li, N, l);
    s.ui = lane_view(b.ui, N, l);
    s.prow = lane_view(b.prow, N, l);
    s.void_slots = lane_view(b.void_slots, N, l);
    s.flag = lane_view(b.flag, N, l);
    s.topo = lane_view(b.topo, N, l);
    s.stack = lane_view(b.stack, N, l);
    s.pstack = lane_view(b.pstack, N, l);
    s.want = lane_view(b.want, N, l);
    s.want_of = lane_view(b.want_of, N, l);
    s.prev = lane_view(b.prev, N, l);
    s.void_col = lane_view(b.void_col, N, l);
    s.scaled_pivot = lane_view(b.scaled_pivot, N, l);
    s.lx = lane_view(b.lx, N, l);
    s.ux = lane_view(b.ux, N, l);
    s.udiag = lane_view(b.udiag, N, l);
    s.w = lane_view(b.w, N, l);
    s.y = lane_view(b.y, N, l);
    s.rscale = lane_view(b.rscale, N, l);
    s.vcopy = lane_view(b.vcopy, N, l);
    LV<double> tri = lane_view(b.tri, N, l);
    s.bp = tri;
    s.mul = tri.at(n);
    s.cv = tri.at(2 * (size_t)n);
    LV<double> pb = lane_view(b.probe, N, l);
    s.px = pb;
    s.pr = pb.at(n);
    s.prows = pb.at(2 * (size_t)n);
    return s;
}

enum { ST_FACTORED = 0, ST_TRI = 1, ST_LU_FACTORED = 2 };
enum { C_NL = 0, C_NU = 1, C_NVOID = 2, C_BASE = 3 };

// ------------------------------------------------------------- Thomas

template <class A>
ESP_HD int tri_factor(const Pattern &P, LaneLU &s, const A &vals)
{
    const uint32_t n = P.n;
    const uint32_t *ap = P.tri_pos, *bpos = P.tri_pos + n, *cpos = P.tri_pos + 2 * n;
    auto tv = [&](uint32_t pos) { return pos != ESP_NONE ? vals[pos] : 0.0; };
    s.bp[0] = tv(bpos[0]);
    s.cv[0] = tv(cpos[0]);
    for (uint32_t i = 0; i < n; i++) {
        if (i > 0) {
            double m = tv(ap[i]) / s.bp[i - 1];
            s.mul[i] = m;
            s.bp[i] = tv(bpos[i]) - m * s.cv[i - 1];
            if (i + 1 < n) s.cv[i] = tv(cpos[i]);
        }
        double d = fabs_(s.bp[i]);
        double below = i + 1 < n ? fabs_(tv(ap[i + 1])) : 0;
        if (!(d > 0 && d < inf_() && d >= PIVOT_TOL * below)) return 1;
    }
    return 0;
}

template <class X>
ESP_HD void tri_solve(const Pattern &P, const LaneLU &s, X &x)
{
    const uint32_t n = P.n;
    for (uint32_t i = 1; i < n; i++) x[i] -= s.mul[i] * x[i - 1];
    x[n - 1] /= s.bp[n - 1];
    for (uint32_t i = n - 1; i > 0;) {
        i--;
        x[i] = (x[i] - s.cv[i] * x[i + 1]) / s.bp[i];
    }
}

// ---------------------------------------------------- Gilbert-Peierls

template <class A>
ESP_HD bool void_unknown(const Pattern &P, const LaneLU &s, const A &vals, uint32_t c)
{
    if (s.pinv[c] != ESP_NONE) return false;
    for (uint32_t p = P.col_ptr[c]; p < P.col_ptr[c + 1]; p++)
        if (vals[p] != 0) return false;
    for (uint32_t j = 0; j < P.n; j++) {
        if (j == c) continue;
        for (uint32_t p = P.col_ptr[j]; p < P.col_ptr[j + 1]; p++)
            if (P.row_idx[p] == c && vals[p] != 0) return false;
    }
    return true;
}

// 0 done, 1 singular, 2 passed the fill cap (unfactored).
template <class A>
ESP_HD int factor_pass(const Pattern &P, LaneLU &s, const A &vals, double pivot_tol, uint64_t cap, bool keep_prev,
                       double keep_min)
{
    const uint32_t n = P.n;
    s.state[ST_LU_FACTORED] = 0;
    for (uint32_t i = 0; i < n; i++) {
        s.pinv[i] = ESP_NONE;
        s.flag[i] = 0;
        s.void_col[i] = 0;
        s.scaled_pivot[i] = 0;
        s.rscale[i] = 0;
    }
    uint32_t nl = 0, nu = 0, nvoid = 0;
    bool has_void = false;
    uint64_t mb = 0;
    for (uint32_t p = 0; p < P.nnz; p++) {
        uint64_t b = bits(fabs_(vals[p]));
        if (b > mb) mb = b;
    }
    if (!(from_bits(mb) <= 1.7976931348623157e308)) return 1;
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = P.col_ptr[j]; p < P.col_ptr[j + 1]; p++) {
            double a = fabs_(vals[p]);
            uint32_t r = P.row_idx[p];
            if (a > s.rscale[r]) s.rscale[r] = a;
        }
    for (uint32_t i = 0; i < n; i++) {
        double r = s.rscale[i];
        s.rscale[i] = (r > 0 && isfinite_(r)) ? 1 / r : 1;
    }
    uint64_t nonfinite = 0;
    for (uint32_t i = 0; i < n; i++) s.want[i] = keep_prev ? s.prev[i] : i;
    for (uint32_t c = 0; c < n; c++) s.want_of[s.want[c]] = c;

    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = P.q[k];
        s.lp[k] = nl;
        s.up[k] = nu;
        const uint32_t mark = k + 1;
        uint32_t nt = 0;
        for (uint32_t p = P.col_ptr[c]; p < P.col_ptr[c + 1]; p++) {
            uint32_t r = P.row_idx[p];
            if (s.flag[r] == mark) continue;
            uint32_t sp = 0;
            s.stack[sp] = r;
            s.flag[r] = mark;
            s.pstack[sp] = s.pinv[r] == ESP_NONE ? 0 : s.lp[s.pinv[r]];
            for (;;) {
                r = s.stack[sp];
                uint32_t kc = s.pinv[r];
                uint32_t end = kc == ESP_NONE ? 0 : s.lp[kc + 1];
                uint32_t pos = s.pstack[sp];
                bool descended = false;
                while (pos < end) {
                    uint32_t child = s.li[pos];
                    pos++;
                    if (s.flag[child] != mark) {
                        s.pstack[sp] = pos;
                        s.flag[child] = mark;
                        sp++;
                        s.stack[sp] = child;
                        s.pstack[sp] = s.pinv[child] == ESP_NONE ? 0 : s.lp[s.pinv[child]];
                        descended = true;
                        break;
                    }
                }
                if (descended) continue;
                s.topo[nt++] = r;
                if (sp == 0) break;
                sp--;
            }
        }
        for (uint32_t p = P.col_ptr[c]; p < P.col_ptr[c + 1]; p++) s.w[P.row_idx[p]] = vals[p];
        for (uint32_t idx = nt; idx > 0;) {
            idx--;
            uint32_t r = s.topo[idx];
            uint32_t kc = s.pinv[r];
            if (kc == ESP_NONE) continue;
            double ukr = s.w[r];
            s.ui[nu] = kc;
            s.ux[nu] = ukr;
            nu++;
            for (uint32_t q = s.lp[kc]; q < s.lp[kc + 1]; q++) s.w[s.li[q]] -= s.lx[q] * ukr;
        }
        double amax = 0, smax = 0;
        uint32_t piv = ESP_NONE;
        for (uint32_t t = 0; t < nt; t++) {
            uint32_t r = s.topo[t];
            if (s.pinv[r] != ESP_NONE) continue;
            double a = fabs_(s.w[r]);
            if (a > amax) {
                amax = a;
                piv = r;
            }
            double sa = a * s.rscale[r];
            if (sa > smax) smax = sa;
        }
        if (piv == ESP_NONE || amax == 0 || !isfinite_(amax)) {
            if (!void_unknown(P, s, vals, c)) {
                for (uint32_t t = 0; t < nt; t++) s.w[s.topo[t]] = 0;
                return 1;
            }
            s.void_col[k] = 1;
            has_void = true;
            s.udiag[k] = 1;
            s.pinv[c] = k;
            for (uint32_t t = 0; t < nt; t++) s.w[s.topo[t]] = 0;
            continue;
        }
        uint32_t pr = s.want[c];
        if (s.pinv[pr] == ESP_NONE) {
            double dmag = fabs_(s.w[pr]);
            if (keep_prev) {
                if (dmag > 0 && dmag >= keep_min * amax) piv = pr;
            } else if (dmag > 0 && dmag >= pivot_tol * amax) {
                piv = pr;
            } else if (dmag > 0 && dmag * s.rscale[pr] >= pivot_tol * smax) {
                piv = pr;
                s.scaled_pivot[k] = 1;
            }
        }
        if (piv != pr) {
            uint32_t c2 = s.want_of[piv];
            s.want[c2] = pr;
            s.want_of[pr] = c2;
        }
        double d = s.w[piv];
        s.udiag[k] = d;
        s.pinv[piv] = k;
        for (uint32_t t = 0; t < nt; t++) {
            uint32_t r = s.topo[t];
            double v = s.w[r];
            nonfinite |= bits(v - v);
            if (s.pinv[r] == ESP_NONE) {
                double l = v / d;
                nonfinite |= bits(l - l);
                s.li[nl] = r;
                s.lx[nl] = l;
                nl++;
            }
            s.w[r] = 0;
        }
        if ((uint64_t)nl + nu > cap) return 2;
    }
    s.lp[n] = nl;
    s.up[n] = nu;
    if (nonfinite != 0) return 1;
    for (uint32_t t = 0; t < nl; t++) s.li[t] = s.pinv[s.li[t]];
    for (uint32_t p = 0; p < P.nnz; p++) s.prow[p] = s.pinv[P.row_idx[p]];
    if (has_void) {
        for (uint32_t j = 0; j < n; j++) {
            uint32_t c = P.q[j];
            for (uint32_t p = P.col_ptr[c]; p < P.col_ptr[c + 1]; p++)
                if (s.void_col[j] || s.void_col[s.prow[p]]) s.void_slots[nvoid++] = p;
        }
    }
    s.cnt[C_NL] = nl;
    s.cnt[C_NU] = nu;
    s.cnt[C_NVOID] = nvoid;
    s.state[ST_LU_FACTORED] = 1;
    return 0;
}

// SparseLu.factor (sparse_lu.zig:274-311) with direct's caps: fill_cap 3,
// keep_min = the growth limit, diag_fill from the template.
template <class A>
ESP_HD int lu_factor(const Pattern &P, LaneLU &s, const A &vals)
{
    const uint32_t n = P.n;
    const double fill_cap = 3.0;
    bool repivot = s.state[ST_LU_FACTORED] && s.cnt[C_BASE] > 0;
    uint64_t cap = P.cap;
    if (repivot) {
        for (uint32_t r = 0; r < n; r++) s.prev[P.q[s.pinv[r]]] = r;
        double c = fill_cap * (double)s.cnt[C_BASE];
        uint64_t cc = (uint64_t)(c < 1e18 ? c : 1e18);
        if (cc < cap) cap = cc;
    } else if (P.diag_fill > 0) {
        for (uint32_t c = 0; c < n; c++) s.prev[c] = c;
        double c = fill_cap * (double)P.diag_fill;
        uint64_t cc = (uint64_t)(c < 1e18 ? c : 1e18);
        if (cc < cap) cap = cc;
    }
    int r = factor_pass(P, s, vals, PIVOT_TOL, cap, false, GROWTH_LIMIT);
    if (r == 1) return 1;
    if (r == 2) {
        for (uint32_t i = 0; i < n; i++) s.w[i] = 0;
        int r2 = factor_pass(P, s, vals, PIVOT_TOL, repivot ? (uint64_t)P.cap : cap, true, GROWTH_LIMIT);
        if (r2 != 0) {
            for (uint32_t i = 0; i < n; i++) s.w[i] = 0;
            return 1;
        }
    }
    if (!repivot) s.cnt[C_BASE] = s.cnt[C_NL] + s.cnt[C_NU];
    return 0;
}

// Replay.column over the lane's own pattern (sparse_lu.zig:1066-1159).
template <class A>
ESP_HD int lu_refactor(const Pattern &P, LaneLU &s, const A &vals, double growth_limit)
{
    const uint32_t n = P.n, nvoid = s.cnt[C_NVOID];
    for (uint32_t t = 0; t < nvoid; t++)
        if (vals[s.void_slots[t]] != 0) return 1;
    uint64_t nonfinite = 0;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t c = P.q[k];
        uint32_t uk0 = s.up[k], uk1 = s.up[k + 1], lk0 = s.lp[k], lk1 = s.lp[k + 1];
        for (uint32_t p = P.col_ptr[c]; p < P.col_ptr[c + 1]; p++) s.w[s.prow[p]] = vals[p];
        for (uint32_t p = uk0; p < uk1; p++) {
            uint32_t i = s.ui[p];
            double uki = s.w[i];
            s.w[i] = 0;
            s.ux[p] = uki;
            nonfinite |= bits(uki - uki);
            for (uint32_t q = s.lp[i]; q < s.lp[i + 1]; q++) s.w[s.li[q]] -= s.lx[q] * uki;
        }
        if (s.void_col[k]) {
            s.udiag[k] = 1;
            for (uint32_t p = lk0; p < lk1; p++) {
                s.lx[p] = 0;
                s.w[s.li[p]] = 0;
            }
            s.w[k] = 0;
            for (uint32_t p = uk0; p < uk1; p++)
                for (uint32_t q = s.lp[s.ui[p]]; q < s.lp[s.ui[p] + 1]; q++) s.w[s.li[q]] = 0;
            continue;
        }
        double d = s.w[k];
        s.w[k] = 0;
        if (d == 0 || !isfinite_(d)) {
            for (uint32_t p = lk0; p < lk1; p++) s.w[s.li[p]] = 0;
            return 1;
        }
        s.udiag[k] = d;
        if (growth_limit > 0 && !s.scaled_pivot[k]) {
            uint64_t cbits = bits(fabs_(d));
            for (uint32_t p = lk0; p < lk1; p++) {
                uint32_t row = s.li[p];
                double v = s.w[row];
                s.w[row] = 0;
                uint64_t b = bits(fabs_(v));
                if ((int64_t)b > (int64_t)cbits) cbits = b;
                s.lx[p] = v / d;
            }
            if (!(fabs_(d) >= growth_limit * from_bits(cbits))) return 1;
            continue;
        }
        for (uint32_t p = lk0; p < lk1; p++) {
            uint32_t row = s.li[p];
            double l = s.w[row] / d;
            nonfinite |= bits(l - l);
            s.lx[p] = l;
            s.w[row] = 0;
        }
    }
    return nonfinite != 0 ? 1 : 0;
}

template <class X, class Y>
ESP_HD void lu_solve(const Pattern &P, const LaneLU &s, const X &b, Y &x)
{
    const uint32_t n = P.n;
    for (uint32_t r = 0; r < n; r++) s.y[s.pinv[r]] = b[r];
    for (uint32_t k = 0; k < n; k++) {
        double yk = s.y[k];
        if (yk == 0) continue;
        for (uint32_t p = s.lp[k]; p < s.lp[k + 1]; p++) s.y[s.li[p]] -= s.lx[p] * yk;
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        double zk = s.y[k] / s.udiag[k];
        s.y[k] = zk;
        if (zk == 0) continue;
        for (uint32_t p = s.up[k]; p < s.up[k + 1]; p++) s.y[s.ui[p]] -= s.ux[p] * zk;
    }
    for (uint32_t j = 0; j < n; j++) x[P.q[j]] = s.y[j];
}

template <class X, class Y>
ESP_HD void lu_solve_t(const Pattern &P, const LaneLU &s, const X &b, Y &x)
{
    const uint32_t n = P.n;
    for (uint32_t j = 0; j < n; j++) s.y[j] = b[P.q[j]];
    for (uint32_t k = 0; k < n; k++) {
        double v = s.y[k];
        for (uint32_t p = s.up[k]; p < s.up[k + 1]; p++) v -= s.ux[p] * s.y[s.ui[p]];
        s.y[k] = v / s.udiag[k];
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        double v = s.y[k];
        for (uint32_t p = s.lp[k]; p < s.lp[k + 1]; p++) v -= s.lx[p] * s.y[s.li[p]];
        s.y[k] = v;
    }
    for (uint32_t r = 0; r < n; r++) x[r] = s.y[s.pinv[r]];
}

template <class X>
ESP_HD double norm_inf_lv(const X &x, uint32_t off, uint32_t len)
{
    uint64_t m = 0;
    for (uint32_t i = 0; i < len; i++) {
        uint64_t b = bits(fabs_(x[off + i]));
        if (b > m) m = b;
    }
    return from_bits(m);
}

ESP_HD double max_multiplier(const Pattern &P, const LaneLU &s)
{
    const uint32_t n = P.n;
    double all = norm_inf_lv(s.lx, 0, s.cnt[C_NL]);
    bool any_scaled = false;
    for (uint32_t k = 0; k < n; k++) any_scaled = any_scaled || s.scaled_pivot[k];
    if (all != all || !any_scaled) return all;
    double m = 0;
    for (uint32_t k = 0; k < n; k++) {
        if (s.scaled_pivot[k]) continue;
        double c = norm_inf_lv(s.lx, s.lp[k], s.lp[k + 1] - s.lp[k]);
        if (c > m) m = c;
    }
    return m;
}

ESP_HD void probe_rhs(LV<double> b, uint32_t n)
{
    double den = (double)((n > 2 ? n : 2) - 1);
    for (uint32_t i = 0; i < n; i++) {
        double mag = 1 + (double)i / den;
        b[i] = i % 2 == 0 ? mag : -mag;
    }
}

template <class A>
ESP_HD bool replay_holds(const Pattern &P, LaneLU &s, const A &vals)
{
    const uint32_t n = P.n;
    if (max_multiplier(P, s) <= 1 / PIVOT_TOL) return true;
    for (uint32_t i = 0; i < n; i++) s.prows[i] = 0;
    double norm_1 = 0;
    for (uint32_t j = 0; j < n; j++) {
        double col = 0;
        for (uint32_t p = P.col_ptr[j]; p < P.col_ptr[j + 1]; p++) {
            double a = fabs_(vals[p]);
            s.prows[P.row_idx[p]] += a;
            col += a;
        }
        if (col > norm_1) norm_1 = col;
    }
    double norm_inf_a = norm_inf_lv(s.prows, 0, n);
    double nf = (double)n;
    for (int transpose = 0; transpose < 2; transpose++) {
        probe_rhs(s.pr, n);
        if (transpose) lu_solve_t(P, s, s.pr, s.px);
        else lu_solve(P, s, s.pr, s.px);
        probe_rhs(s.pr, n);
        for (uint32_t j = 0; j < n; j++)
            for (uint32_t p = P.col_ptr[j]; p < P.col_ptr[j + 1]; p++) {
                uint32_t i = P.row_idx[p];
                if (transpose) s.pr[j] -= vals[p] * s.px[i];
                else s.pr[i] -= vals[p] * s.px[j];
            }
        double a_norm = transpose ? norm_1 : norm_inf_a;
        double ratio = norm_inf_lv(s.pr, 0, n) / (a_norm * norm_inf_lv(s.px, 0, n) * nf * EPS);
        if (!(ratio <= 30.0 / PIVOT_TOL)) return false;
    }
    return true;
}

// direct.Solver.factor for one lane; 0 factored, 1 singular.
template <class A>
ESP_HD int direct_factor(const Pattern &P, LaneLU &s, const A &vals)
{
    const uint32_t nnz = P.nnz;
    if (nnz != 0 && s.state[ST_FACTORED]) {
        bool eq = true;
        for (uint32_t p = 0; p < nnz; p++) eq = eq && vals[p] == s.vcopy[p];
        if (eq) return 0;
    }
    if (s.state[ST_TRI]) {
        if (tri_factor(P, s, vals) == 0) {
            s.state[ST_FACTORED] = 1;
            for (uint32_t p = 0; p < nnz; p++) s.vcopy[p] = vals[p];
            return 0;
        }
        s.state[ST_TRI] = 0;
        s.state[ST_FACTORED] = 0;
    }
    if (s.state[ST_FACTORED]) {
        if (lu_refactor(P, s, vals, GROWTH_LIMIT) == 0 && replay_holds(P, s, vals)) {
            for (uint32_t p = 0; p < nnz; p++) s.vcopy[p] = vals[p];
            return 0;
        }
        s.state[ST_FACTORED] = 0;
    }
    if (lu_factor(P, s, vals) != 0) return 1;
    s.state[ST_FACTORED] = 1;
    for (uint32_t p = 0; p < nnz; p++) s.vcopy[p] = vals[p];
    return 0;
}

// dx = -A^-1 rhs after direct_factor (direct.zig:574-579).
template <class R, class D>
ESP_HD void direct_solve_neg(const Pattern &P, LaneLU &s, const R &rhs, D &dx)
{
    const uint32_t n = P.n;
    if (s.state[ST_TRI]) {
        for (uint32_t i = 0; i < n; i++) dx[i] = -1.0 * rhs[i];
        tri_solve(P, s, dx);
        return;
    }
    for (uint32_t i = 0; i < n; i++) s.px[i] = -1.0 * rhs[i];
    lu_solve(P, s, s.px, dx);
}

// ------------------------------------------------- lead-tape lane bodies

// `LaneTab`, src/solver/kernels/lane.zig:17-37 (word offsets into idx).
struct LaneTab {
    uint32_t n, nnz, nnz_l, nnz_u;
    double growth, wide;
    uint32_t col_ptr, prow, q, pinv, lp, li, up, ui, flags, void_slots, n_void_slots, _pad;
};

// lane.zig:56-125 `Bodies.refactor` (+ wideMultipliers): true when the lane
// must peel. `a` is lane-major in pattern order; `w` is zero on entry/exit.
template <class A>
ESP_HD bool tab_refactor(const LaneTab &t, const uint32_t *x, const A &a, LV<double> lx, LV<double> ux, LV<double> udiag,
                         LV<double> w)
{
    bool bad = false;
    uint64_t nonfinite = 0;
    for (uint32_t v = 0; v < t.n_void_slots; v++)
        if (a[x[t.void_slots + v]] != 0) return true;
    for (uint32_t k = 0; k < t.n; k++) {
        uint32_t c = x[t.q + k];
        for (uint32_t p = x[t.col_ptr + c]; p < x[t.col_ptr + c + 1]; p++) w[x[t.prow + p]] = a[p];
        for (uint32_t p = x[t.up + k]; p < x[t.up + k + 1]; p++) {
            uint32_t i = x[t.ui + p];
            double uki = w[i];
            w[i] = 0;
            ux[p] = uki;
            nonfinite |= bits(uki - uki);
            for (uint32_t pl = x[t.lp + i]; pl < x[t.lp + i + 1]; pl++) w[x[t.li + pl]] -= lx[pl] * uki;
        }
        if (x[t.flags + k] & 1u) {
            udiag[k] = 1;
            for (uint32_t p = x[t.lp + k]; p < x[t.lp + k + 1]; p++) {
                lx[p] = 0;
                w[x[t.li + p]] = 0;
            }
            w[k] = 0;
            for (uint32_t p = x[t.up + k]; p < x[t.up + k + 1]; p++) {
                uint32_t i = x[t.ui + p];
                for (uint32_t pl = x[t.lp + i]; pl < x[t.lp + i + 1]; pl++) w[x[t.li + pl]] = 0;
            }
            continue;
        }
        double d = w[k];
        w[k] = 0;
        bool dead = d == 0 || !(fabs_(d) < inf_());
        bad = bad || dead;
        if (dead) d = 1;
        udiag[k] = d;
        if (t.growth > 0 && (x[t.flags + k] & 2u) == 0) {
            int64_t cbits = (int64_t)bits(fabs_(d));
            for (uint32_t p = x[t.lp + k]; p < x[t.lp + k + 1]; p++) {
                uint32_t r = x[t.li + p];
                double v = w[r];
                w[r] = 0;
                int64_t b = (int64_t)bits(fabs_(v));
                if (b > cbits) cbits = b;
                lx[p] = v / d;
            }
            bad = bad || !(fabs_(d) >= t.growth * from_bits((uint64_t)cbits));
        } else {
            for (uint32_t p = x[t.lp + k]; p < x[t.lp + k + 1]; p++) {
                uint32_t r = x[t.li + p];
                double lv = w[r] / d;
                nonfinite |= bits(lv - lv);
                lx[p] = lv;
                w[r] = 0;
            }
        }
    }
    if (nonfinite != 0) return true;
    uint64_t worst = 0;
    for (uint32_t k = 0; k < t.n; k++) {
        if (x[t.flags + k] & 2u) continue;
        for (uint32_t p = x[t.lp + k]; p < x[t.lp + k + 1]; p++) {
            uint64_t b = bits(fabs_(lx[p]));
            if (b > worst) worst = b;
        }
    }
    return bad || !(from_bits(worst) <= t.wide);
}

// lane.zig:130-157 `Bodies.solveNeg`.
template <class R, class D>
ESP_HD void tab_solve_neg(const LaneTab &t, const uint32_t *x, LV<double> lx, LV<double> ux, LV<double> udiag, const R &rhs,
                          LV<double> y, D &dx)
{
    for (uint32_t r = 0; r < t.n; r++) y[x[t.pinv + r]] = -rhs[r];
    for (uint32_t k = 0; k < t.n; k++) {
        double yk = y[k];
        if (yk != 0)
            for (uint32_t p = x[t.lp + k]; p < x[t.lp + k + 1]; p++) y[x[t.li + p]] -= lx[p] * yk;
    }
    for (uint32_t k = t.n; k > 0;) {
        k--;
        double zk = y[k] / udiag[k];
        y[k] = zk;
        if (zk != 0)
            for (uint32_t p = x[t.up + k]; p < x[t.up + k + 1]; p++) y[x[t.ui + p]] -= ux[p] * zk;
    }
    for (uint32_t j = 0; j < t.n; j++) dx[x[t.q + j]] = y[j];
}

}  // namespace esp
