/* Direct solver facade oracle: restates src/solver/direct.zig for the
 * sparse-LU engine. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   or_is_tridiag       tridiag.zig:15-26 (the reference then runs the
 *                       Thomas engine, which has no lane tape; GPU lanes
 *                       refuse such a pattern, lanes.zig:595-605)
 *   or_solver_new       direct.zig:147-172, initLu :218-240 (BTF+AMD order,
 *                       diag_fill = max(diagFill, nnz)); _ordered shares
 *                       one order across the instances of a pattern
 *   or_solver_factor    direct.zig:310-320, factorInner :348-383 (unchanged
 *                       bypass, refactor + replayHolds, else full factor;
 *                       fill_cap 3, keep_min = growth limit 1e-12)
 *   replay_holds        direct.zig:402-447, maxMultiplier :674-697,
 *                       probeRhs :702-708, replay_ratio_max 30 (:667)
 *   or_solver_solve_neg direct.zig:574-579
 *   tri_*               tridiag.zig:57-137 TriDiag (Thomas; a pivot below
 *                       pivot_tol times the entry under it fails and
 *                       direct.zig:359-362 demotes to the LU for good)
 * BBD (>= 8 subcircuit blocks) is not restated: the lane templates here
 * carry no subcircuit partition, so the reference would pick the LU.
 * Parity pinned by: tests/test_oracle_lu.py (bypass, repivot on a failed
 * replay, probe) and through the lane Newton tests.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int or_lu_sn_reached(const or_lu *lu);

struct or_solver {
    int tri; /* live engine is the Thomas solver */
    uint32_t *a_pos, *b_pos, *c_pos;
    double *bp, *mul, *cv;
    uint32_t n;
    const uint32_t *col_ptr, *row_idx;
    uint32_t *q;
    or_lu *lu;
    double *vcopy;
    double *probe;
    uint64_t diag_fill;
    int factored;
};

int or_is_tridiag(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx)
{
    if (n < 3) return 0;
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            uint32_t r = row_idx[p];
            uint32_t diff = r >= j ? r - j : j - r;
            if (diff > 1) return 0;
        }
    return 1;
}

static double tri_val(uint32_t pos, const double *vals) { return pos != OR_NONE ? vals[pos] : 0; }

static int tri_factor(or_solver *s, const double *vals)
{
    const uint32_t n = s->n;
    s->bp[0] = tri_val(s->b_pos[0], vals);
    s->cv[0] = tri_val(s->c_pos[0], vals);
    for (uint32_t i = 0; i < n; i++) {
        if (i > 0) {
            double m = tri_val(s->a_pos[i], vals) / s->bp[i - 1];
            s->mul[i] = m;
            s->bp[i] = tri_val(s->b_pos[i], vals) - m * s->cv[i - 1];
            if (i + 1 < n) s->cv[i] = tri_val(s->c_pos[i], vals);
        }
        double d = fabs(s->bp[i]);
        double below = i + 1 < n ? fabs(tri_val(s->a_pos[i + 1], vals)) : 0;
        if (!(d > 0 && d < INFINITY && d >= OR_PIVOT_TOL * below)) return 1;
    }
    return 0;
}

static void tri_solve(const or_solver *s, double *x)
{
    const uint32_t n = s->n;
    for (uint32_t i = 1; i < n; i++) x[i] -= s->mul[i] * x[i - 1];
    x[n - 1] /= s->bp[n - 1];
    for (uint32_t i = n - 1; i > 0;) {
        i--;
        x[i] = (x[i] - s->cv[i] * x[i + 1]) / s->bp[i];
    }
}

int or_solver_is_tri(const or_solver *s) { return s->tri; }

/* A solver on a column order computed once for the pattern and shared by
 * every instance of it (device/Circuit.zig:55-60 ColumnOrder, direct.zig:
 * 143-152 initOrdered): `q` is copied, `diag_fill` is or_diag_fill's count. */
or_solver *or_solver_new_ordered(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const uint32_t *q,
                                 uint64_t diag_fill)
{
    or_solver *s = calloc(1, sizeof(or_solver));
    if (or_is_tridiag(n, col_ptr, row_idx)) {
        s->tri = 1;
        s->a_pos = malloc(4 * 3 * (size_t)n);
        s->b_pos = s->a_pos + n;
        s->c_pos = s->a_pos + 2 * n;
        for (uint32_t i = 0; i < 3 * n; i++) s->a_pos[i] = OR_NONE;
        s->bp = calloc(3 * (size_t)n, 8);
        s->mul = s->bp + n;
        s->cv = s->bp + 2 * n;
        for (uint32_t j = 0; j < n; j++)
            for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
                uint32_t r = row_idx[p];
                if (r == j) s->b_pos[j] = p;
                else if (r == j + 1) s->a_pos[j + 1] = p;
                else if (j > 0 && r + 1 == j) s->c_pos[r] = p;
            }
    }
    s->n = n;
    s->col_ptr = col_ptr;
    s->row_idx = row_idx;
    s->q = malloc(sizeof(uint32_t) * (n ? n : 1));
    if (n) memcpy(s->q, q, sizeof(uint32_t) * n);
    uint64_t df = diag_fill;
    if (df < col_ptr[n]) df = col_ptr[n];
    s->diag_fill = df;
    s->lu = or_lu_new(n, col_ptr, s->q);
    or_lu_set_caps(s->lu, 0, 0, df);
    s->vcopy = malloc(sizeof(double) * (col_ptr[n] ? col_ptr[n] : 1));
    return s;
}

or_solver *or_solver_new(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx)
{
    uint32_t *q = malloc(sizeof(uint32_t) * (n ? n : 1));
    if (or_order(n, col_ptr, row_idx, q) != 0) {
        free(q);
        return NULL;
    }
    or_solver *s = or_solver_new_ordered(n, col_ptr, row_idx, q, or_diag_fill(n, col_ptr, row_idx, q));
    free(q);
    return s;
}

void or_solver_free(or_solver *s)
{
    if (!s) return;
    or_lu_free(s->lu);
    free(s->q);
    free(s->vcopy);
    free(s->probe);
    free(s->a_pos);
    free(s->bp);
    free(s);
}

or_lu *or_solver_lu(or_solver *s) { return s->lu; }
const uint32_t *or_solver_q(const or_solver *s) { return s->q; }

static double norm_inf(const double *x, size_t len)
{
    uint64_t m = 0;
    for (size_t i = 0; i < len; i++) {
        double a = fabs(x[i]);
        uint64_t b;
        memcpy(&b, &a, 8);
        if (b > m) m = b;
    }
    double r;
    memcpy(&r, &m, 8);
    return r;
}

static double max_multiplier(const or_lu *lu)
{
    const uint32_t n = or_lu_n(lu);
    const double *lx = or_lu_lx(lu);
    const uint32_t *lp = or_lu_lp(lu);
    const uint8_t *sp = or_lu_scaled_pivot(lu);
    double all = norm_inf(lx, or_lu_nnz_l(lu));
    int any_scaled = 0;
    for (uint32_t k = 0; k < n; k++) any_scaled |= sp[k];
    if (isnan(all) || !any_scaled) return all;
    double m = 0;
    for (uint32_t k = 0; k < n; k++) {
        if (sp[k]) continue;
        double c = norm_inf(lx + lp[k], lp[k + 1] - lp[k]);
        if (c > m) m = c;
    }
    return m;
}

static void probe_rhs(double *b, uint32_t n)
{
    double den = (double)((n > 2 ? n : 2) - 1);
    for (uint32_t i = 0; i < n; i++) {
        double mag = 1 + (double)i / den;
        b[i] = i % 2 == 0 ? mag : -mag;
    }
}

static int replay_holds(or_solver *s, const double *vals)
{
    const uint32_t n = s->n;
    if (max_multiplier(s->lu) <= 1 / OR_PIVOT_TOL) return 1;
    if (!s->probe) s->probe = malloc(sizeof(double) * 3 * (n ? n : 1));
    double *x = s->probe, *r = s->probe + n, *rows = s->probe + 2 * n;
    memset(rows, 0, sizeof(double) * n);
    double norm_1 = 0;
    for (uint32_t j = 0; j < n; j++) {
        double col = 0;
        for (uint32_t p = s->col_ptr[j]; p < s->col_ptr[j + 1]; p++) {
            double a = fabs(vals[p]);
            rows[s->row_idx[p]] += a;
            col += a;
        }
        if (col > norm_1) norm_1 = col;
    }
    double norm_inf_a = norm_inf(rows, n);
    double nf = (double)n;
    for (int transpose = 0; transpose < 2; transpose++) {
        probe_rhs(r, n);
        if (transpose) or_lu_solve_t(s->lu, r, x);
        else or_lu_solve(s->lu, r, x);
        probe_rhs(r, n);
        for (uint32_t j = 0; j < n; j++)
            for (uint32_t p = s->col_ptr[j]; p < s->col_ptr[j + 1]; p++) {
                uint32_t i = s->row_idx[p];
                if (transpose) r[j] -= vals[p] * x[i];
                else r[i] -= vals[p] * x[j];
            }
        double a_norm = transpose ? norm_1 : norm_inf_a;
        double ratio = norm_inf(r, n) / (a_norm * norm_inf(x, n) * nf * 2.220446049250313e-16);
        if (!(ratio <= 30.0 / OR_PIVOT_TOL)) return 0;
    }
    return 1;
}

int or_solver_factor(or_solver *s, const double *vals)
{
    const uint32_t nnz = s->col_ptr[s->n];
    if (nnz != 0 && s->factored) {
        int eq = 1;
        for (uint32_t p = 0; p < nnz; p++) eq &= vals[p] == s->vcopy[p];
        if (eq) return 0;
    }
    if (s->tri) {
        if (tri_factor(s, vals) == 0) {
            s->factored = 1;
            memcpy(s->vcopy, vals, sizeof(double) * nnz);
            return 0;
        }
        s->tri = 0; /* demote */
        s->factored = 0;
    }
    or_lu_set_caps(s->lu, 3.0, OR_GROWTH_LIMIT, s->diag_fill);
    if (s->factored) {
        if (or_lu_refactor(s->lu, s->col_ptr, vals, OR_GROWTH_LIMIT) == 0 && replay_holds(s, vals)) {
            memcpy(s->vcopy, vals, sizeof(double) * nnz);
            return 0;
        }
        s->factored = 0;
    }
    if (or_lu_factor(s->lu, s->col_ptr, s->row_idx, vals, OR_PIVOT_TOL) != 0) return 1;
    s->factored = 1;
    memcpy(s->vcopy, vals, sizeof(double) * nnz);
    return 0;
}

void or_solver_solve_neg(or_solver *s, const double *rhs, double *dx)
{
    if (s->tri) {
        for (uint32_t i = 0; i < s->n; i++) dx[i] = -1.0 * rhs[i];
        tri_solve(s, dx);
        return;
    }
    double *b = malloc(sizeof(double) * (s->n ? s->n : 1));
    for (uint32_t i = 0; i < s->n; i++) b[i] = -1.0 * rhs[i];
    or_lu_solve(s->lu, b, dx);
    free(b);
}
