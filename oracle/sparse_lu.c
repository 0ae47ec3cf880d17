/* Sparse LU oracle: restates src/solver/sparse_lu.zig. TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 *
 *   or_lu_factor    sparse_lu.zig:274-311  factor with the fill caps
 *   factor_pass     sparse_lu.zig:337-663  left-looking Gilbert-Peierls:
 *                   DFS reach of A[:,c] through L, triangular solve in
 *                   reverse finish order, threshold pivot (diagonal first,
 *                   row-scaled second chance, want hand-off), void unknown
 *                   unit pivot, L/U store with the non-finite flag
 *   void_unknown    sparse_lu.zig:916-935
 *   or_lu_refactor  sparse_lu.zig:973-1011 + Replay.column :1066-1159
 *                   (the column replay; the flat tape of :1177-1271 is its
 *                   bitwise equal and is not restated)
 *   or_lu_solve     sparse_lu.zig:1277-1321
 *   or_lu_solve_t   sparse_lu.zig:1326-... (A^T x = b)
 * The plain DFS walk is restated. The reference switches to a supernodal
 * copy of the loop once an L column reaches 32 rows (panel_min_rows,
 * :42); that copy keeps the plain walk's order until panels hold 4/5 of L
 * (`sn_dfs`, :617), after which the walk order (and so U's row order) can
 * differ. `or_lu_sn_reached` reports when a factor crossed 32 rows, so a
 * caller knows its tape is outside the restated path: parity unpinned for
 * such (mesh-sized) factors.
 * Parity pinned by: tests/test_oracle_lu.py (sparse_lu.zig:1450-1527 scaled
 * residual bars and refactor == factor bitwise on random systems;
 * :1529-1547 void column with NaN replay; :1552-1583 overflow case;
 * :1622-1660 fill cap restart; :1662-1680 empty and 1x1; :1682-1740 power
 * of two scaling exactness; LAPACK/LAWN 41 residuals).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NONE OR_NONE
#define PANEL_MIN_ROWS 32

typedef struct { uint32_t *p; size_t len, cap; } vec_u32;
typedef struct { double *p; size_t len, cap; } vec_f64;

static void vu_push(vec_u32 *v, uint32_t x)
{
    if (v->len == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 64;
        v->p = realloc(v->p, sizeof(uint32_t) * v->cap);
    }
    v->p[v->len++] = x;
}

static void vf_push(vec_f64 *v, double x)
{
    if (v->len == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 64;
        v->p = realloc(v->p, sizeof(double) * v->cap);
    }
    v->p[v->len++] = x;
}

struct or_lu {
    uint32_t n, nnz;
    uint32_t *q; /* owned copy */
    uint32_t *pinv, *lp, *up, *prow;
    vec_u32 li, ui, void_slots;
    vec_f64 lx, ux;
    double *udiag, *w, *y, *rscale;
    uint8_t *void_col, *scaled_pivot;
    uint32_t *want, *want_of, *prev, *flag, *topo, *stack, *pstack;
    size_t base_lu;
    uint64_t diag_fill;
    double fill_cap, keep_min;
    int factored;
    int sn_reached;
};

or_lu *or_lu_new(uint32_t n, const uint32_t *col_ptr, const uint32_t *q)
{
    or_lu *lu = calloc(1, sizeof(or_lu));
    size_t m = n ? n : 1;
    lu->n = n;
    lu->nnz = col_ptr[n];
    lu->q = malloc(sizeof(uint32_t) * m);
    if (n) memcpy(lu->q, q, sizeof(uint32_t) * n);
    lu->pinv = malloc(sizeof(uint32_t) * m);
    lu->lp = calloc(n + 1, sizeof(uint32_t));
    lu->up = calloc(n + 1, sizeof(uint32_t));
    lu->prow = malloc(sizeof(uint32_t) * (lu->nnz ? lu->nnz : 1));
    lu->udiag = calloc(m, sizeof(double));
    lu->w = calloc(m, sizeof(double));
    lu->y = calloc(m, sizeof(double));
    lu->rscale = calloc(m, sizeof(double));
    lu->void_col = calloc(m, 1);
    lu->scaled_pivot = calloc(m, 1);
    lu->want = malloc(sizeof(uint32_t) * m);
    lu->want_of = malloc(sizeof(uint32_t) * m);
    lu->prev = malloc(sizeof(uint32_t) * m);
    lu->flag = malloc(sizeof(uint32_t) * m);
    lu->topo = malloc(sizeof(uint32_t) * m);
    lu->stack = malloc(sizeof(uint32_t) * m);
    lu->pstack = malloc(sizeof(uint32_t) * m);
    return lu;
}

void or_lu_free(or_lu *lu)
{
    if (!lu) return;
    free(lu->q); free(lu->pinv); free(lu->lp); free(lu->up); free(lu->prow);
    free(lu->li.p); free(lu->ui.p); free(lu->void_slots.p); free(lu->lx.p); free(lu->ux.p);
    free(lu->udiag); free(lu->w); free(lu->y); free(lu->rscale); free(lu->void_col);
    free(lu->scaled_pivot); free(lu->want); free(lu->want_of); free(lu->prev); free(lu->flag);
    free(lu->topo); free(lu->stack); free(lu->pstack);
    free(lu);
}

void or_lu_set_caps(or_lu *lu, double fill_cap, double keep_min, uint64_t diag_fill)
{
    lu->fill_cap = fill_cap;
    lu->keep_min = keep_min;
    lu->diag_fill = diag_fill;
}

static int void_unknown(const or_lu *lu, const uint32_t *col_ptr, const uint32_t *row_idx, const double *vals, uint32_t c)
{
    if (lu->pinv[c] != NONE) return 0;
    for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++)
        if (vals[p] != 0) return 0;
    for (uint32_t j = 0; j < lu->n; j++) {
        if (j == c) continue;
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++)
            if (row_idx[p] == c && vals[p] != 0) return 0;
    }
    return 1;
}

static uint64_t bits_of(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* 0 done, 1 singular, 2 passed the fill cap (unfactored). */
static int factor_pass(or_lu *lu, const uint32_t *col_ptr, const uint32_t *row_idx, const double *vals, double pivot_tol, size_t cap, int keep_prev)
{
    const uint32_t n = lu->n;
    uint32_t *pinv = lu->pinv, *flag = lu->flag, *stack = lu->stack, *pstack = lu->pstack;
    uint32_t *topo = lu->topo, *lp = lu->lp, *want = lu->want, *want_of = lu->want_of;
    double *w = lu->w, *rscale = lu->rscale;
    lu->factored = 0;
    lu->sn_reached = 0;
    for (uint32_t i = 0; i < n; i++) {
        pinv[i] = NONE;
        flag[i] = 0;
        lu->void_col[i] = 0;
        lu->scaled_pivot[i] = 0;
        rscale[i] = 0;
    }
    lu->void_slots.len = 0;
    lu->li.len = lu->lx.len = lu->ui.len = lu->ux.len = 0;
    int has_void = 0;

    /* normInf over the pattern values: the max of |v|'s bit patterns, so a
     * NaN (whose pattern sorts above inf) wins. */
    uint64_t mb = 0;
    for (uint32_t p = 0; p < lu->nnz; p++) {
        uint64_t b = bits_of(fabs(vals[p]));
        if (b > mb) mb = b;
    }
    double vmax;
    memcpy(&vmax, &mb, 8);
    if (!(vmax <= 1.7976931348623157e308)) return 1;
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            double a = fabs(vals[p]);
            if (a > rscale[row_idx[p]]) rscale[row_idx[p]] = a;
        }
    for (uint32_t i = 0; i < n; i++)
        rscale[i] = (rscale[i] > 0 && isfinite(rscale[i])) ? 1 / rscale[i] : 1;

    uint64_t nonfinite = 0;
    for (uint32_t i = 0; i < n; i++) want[i] = keep_prev ? lu->prev[i] : i;
    for (uint32_t c = 0; c < n; c++) want_of[want[c]] = c;

    for (uint32_t k = 0; k < n; k++) {
        const uint32_t c = lu->q[k];
        lp[k] = (uint32_t)lu->li.len;
        lu->up[k] = (uint32_t)lu->ui.len;
        const uint32_t mark = k + 1;
        const uint32_t *li = lu->li.p;

        uint32_t nt = 0;
        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) {
            uint32_t r = row_idx[p];
            if (flag[r] == mark) continue;
            uint32_t sp = 0;
            stack[sp] = r;
            flag[r] = mark;
            pstack[sp] = pinv[r] == NONE ? 0 : lp[pinv[r]];
            for (;;) {
                r = stack[sp];
                uint32_t kc = pinv[r];
                uint32_t end = kc == NONE ? 0 : lp[kc + 1];
                uint32_t pos = pstack[sp];
                int descended = 0;
                while (pos < end) {
                    uint32_t child = li[pos];
                    pos++;
                    if (flag[child] != mark) {
                        pstack[sp] = pos;
                        flag[child] = mark;
                        sp++;
                        stack[sp] = child;
                        pstack[sp] = pinv[child] == NONE ? 0 : lp[pinv[child]];
                        descended = 1;
                        break;
                    }
                }
                if (descended) continue;
                topo[nt++] = r;
                if (sp == 0) break;
                sp--;
            }
        }

        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) w[row_idx[p]] = vals[p];

        for (uint32_t idx = nt; idx > 0;) {
            idx--;
            uint32_t r = topo[idx];
            uint32_t kc = pinv[r];
            if (kc == NONE) continue;
            double ukr = w[r];
            vu_push(&lu->ui, kc);
            vf_push(&lu->ux, ukr);
            /* scatterAxpy: dst[idx[p]] -= src[p] * f, two at a time; a
             * column's rows are distinct so the pairing is the plain loop. */
            for (uint32_t q = lp[kc]; q < lp[kc + 1]; q++) w[lu->li.p[q]] -= lu->lx.p[q] * ukr;
        }

        double amax = 0, smax = 0;
        uint32_t piv = NONE;
        for (uint32_t t = 0; t < nt; t++) {
            uint32_t r = topo[t];
            if (pinv[r] != NONE) continue;
            double a = fabs(w[r]);
            if (a > amax) {
                amax = a;
                piv = r;
            }
            double sa = a * rscale[r];
            if (sa > smax) smax = sa;
        }
        if (piv == NONE || amax == 0 || !isfinite(amax)) {
            if (!void_unknown(lu, col_ptr, row_idx, vals, c)) {
                for (uint32_t t = 0; t < nt; t++) w[topo[t]] = 0;
                return 1;
            }
            lu->void_col[k] = 1;
            has_void = 1;
            lu->udiag[k] = 1;
            pinv[c] = k;
            for (uint32_t t = 0; t < nt; t++) w[topo[t]] = 0;
            continue;
        }
        uint32_t pr = want[c];
        if (pinv[pr] == NONE) {
            double dmag = fabs(w[pr]);
            if (keep_prev) {
                if (dmag > 0 && dmag >= lu->keep_min * amax) piv = pr;
            } else if (dmag > 0 && dmag >= pivot_tol * amax) {
                piv = pr;
            } else if (dmag > 0 && dmag * rscale[pr] >= pivot_tol * smax) {
                piv = pr;
                lu->scaled_pivot[k] = 1;
            }
        }
        if (piv != pr) {
            uint32_t c2 = want_of[piv];
            want[c2] = pr;
            want_of[pr] = c2;
        }
        double d = w[piv];
        lu->udiag[k] = d;
        pinv[piv] = k;
        for (uint32_t t = 0; t < nt; t++) {
            uint32_t r = topo[t];
            double v = w[r];
            double vv = v - v;
            nonfinite |= bits_of(vv);
            if (pinv[r] == NONE) {
                double l = v / d;
                double ll = l - l;
                nonfinite |= bits_of(ll);
                vu_push(&lu->li, r);
                vf_push(&lu->lx, l);
            }
            w[r] = 0;
        }
        if (lu->li.len + lu->ui.len > cap) return 2;
        if (lu->li.len - lp[k] >= PANEL_MIN_ROWS) lu->sn_reached = 1;
    }
    lp[n] = (uint32_t)lu->li.len;
    lu->up[n] = (uint32_t)lu->ui.len;
    if (nonfinite != 0) return 1;
    for (size_t t = 0; t < lu->li.len; t++) lu->li.p[t] = pinv[lu->li.p[t]];
    for (uint32_t p = 0; p < lu->nnz; p++) lu->prow[p] = pinv[row_idx[p]];
    if (has_void) {
        for (uint32_t j = 0; j < n; j++) {
            uint32_t c = lu->q[j];
            for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++)
                if (lu->void_col[j] || lu->void_col[lu->prow[p]]) vu_push(&lu->void_slots, p);
        }
    }
    lu->factored = 1;
    return 0;
}

int or_lu_factor(or_lu *lu, const uint32_t *col_ptr, const uint32_t *row_idx, const double *vals, double pivot_tol)
{
    int repivot = lu->factored && lu->fill_cap > 0 && lu->base_lu > 0;
    size_t cap = (size_t)-1;
    if (repivot) {
        for (uint32_t r = 0; r < lu->n; r++) lu->prev[lu->q[lu->pinv[r]]] = r;
        double c = lu->fill_cap * (double)lu->base_lu;
        cap = (size_t)(c < 1e18 ? c : 1e18);
    } else if (lu->fill_cap > 0 && lu->diag_fill > 0) {
        for (uint32_t c = 0; c < lu->n; c++) lu->prev[c] = c;
        double c = lu->fill_cap * (double)lu->diag_fill;
        cap = (size_t)(c < 1e18 ? c : 1e18);
    }
    int r = factor_pass(lu, col_ptr, row_idx, vals, pivot_tol, cap, 0);
    if (r == 1) return 1;
    if (r == 2) {
        for (uint32_t i = 0; i < lu->n; i++) lu->w[i] = 0;
        int r2 = factor_pass(lu, col_ptr, row_idx, vals, pivot_tol, repivot ? (size_t)-1 : cap, 1);
        if (r2 != 0) {
            for (uint32_t i = 0; i < lu->n; i++) lu->w[i] = 0;
            return 1;
        }
    }
    if (!repivot) lu->base_lu = lu->li.len + lu->ui.len;
    return 0;
}

int or_lu_refactor(or_lu *lu, const uint32_t *col_ptr, const double *vals, double growth_limit)
{
    for (size_t t = 0; t < lu->void_slots.len; t++)
        if (vals[lu->void_slots.p[t]] != 0) return 1;
    const uint32_t *li = lu->li.p, *ui = lu->ui.p, *lp = lu->lp;
    double *lx = lu->lx.p, *ux = lu->ux.p, *w = lu->w;
    uint64_t nonfinite = 0;
    for (uint32_t k = 0; k < lu->n; k++) {
        uint32_t c = lu->q[k];
        uint32_t uk0 = lu->up[k], uk1 = lu->up[k + 1], lk0 = lp[k], lk1 = lp[k + 1];
        for (uint32_t p = col_ptr[c]; p < col_ptr[c + 1]; p++) w[lu->prow[p]] = vals[p];
        for (uint32_t p = uk0; p < uk1; p++) {
            uint32_t i = ui[p];
            double uki = w[i];
            w[i] = 0;
            ux[p] = uki;
            double vv = uki - uki;
            nonfinite |= bits_of(vv);
            for (uint32_t q = lp[i]; q < lp[i + 1]; q++) w[li[q]] -= lx[q] * uki;
        }
        if (lu->void_col[k]) {
            lu->udiag[k] = 1;
            for (uint32_t p = lk0; p < lk1; p++) {
                lx[p] = 0;
                w[li[p]] = 0;
            }
            w[k] = 0;
            for (uint32_t p = uk0; p < uk1; p++)
                for (uint32_t q = lp[ui[p]]; q < lp[ui[p] + 1]; q++) w[li[q]] = 0;
            continue;
        }
        double d = w[k];
        w[k] = 0;
        if (d == 0 || !isfinite(d)) {
            for (uint32_t p = lk0; p < lk1; p++) w[li[p]] = 0;
            return 1;
        }
        lu->udiag[k] = d;
        if (growth_limit > 0 && !lu->scaled_pivot[k]) {
            uint64_t cbits = bits_of(fabs(d));
            for (uint32_t p = lk0; p < lk1; p++) {
                uint32_t row = li[p];
                double v = w[row];
                w[row] = 0;
                uint64_t b = bits_of(fabs(v));
                /* i64 max of nonnegative patterns: NaN sorts above inf */
                if ((int64_t)b > (int64_t)cbits) cbits = b;
                lx[p] = v / d;
            }
            double cmax;
            memcpy(&cmax, &cbits, 8);
            if (!(fabs(d) >= growth_limit * cmax)) return 1;
            continue;
        }
        for (uint32_t p = lk0; p < lk1; p++) {
            uint32_t row = li[p];
            double l = w[row] / d;
            double ll = l - l;
            nonfinite |= bits_of(ll);
            lx[p] = l;
            w[row] = 0;
        }
    }
    if (nonfinite != 0) return 1;
    return 0;
}

void or_lu_solve(const or_lu *lu, const double *b, double *x)
{
    const uint32_t n = lu->n;
    double *y = lu->y;
    for (uint32_t r = 0; r < n; r++) y[lu->pinv[r]] = b[r];
    for (uint32_t k = 0; k < n; k++) {
        double yk = y[k];
        if (yk == 0) continue;
        for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) y[lu->li.p[p]] -= lu->lx.p[p] * yk;
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        double zk = y[k] / lu->udiag[k];
        y[k] = zk;
        if (zk == 0) continue;
        for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) y[lu->ui.p[p]] -= lu->ux.p[p] * zk;
    }
    for (uint32_t j = 0; j < n; j++) x[lu->q[j]] = y[j];
}

/* A^T x = b: A = P^-1 L U Q^-1, so solve U^T then L^T. */
void or_lu_solve_t(const or_lu *lu, const double *b, double *x)
{
    const uint32_t n = lu->n;
    double *y = lu->y;
    for (uint32_t j = 0; j < n; j++) y[j] = b[lu->q[j]];
    for (uint32_t k = 0; k < n; k++) {
        double s = y[k];
        for (uint32_t p = lu->up[k]; p < lu->up[k + 1]; p++) s -= lu->ux.p[p] * y[lu->ui.p[p]];
        y[k] = s / lu->udiag[k];
    }
    for (uint32_t k = n; k > 0;) {
        k--;
        double s = y[k];
        for (uint32_t p = lu->lp[k]; p < lu->lp[k + 1]; p++) s -= lu->lx.p[p] * y[lu->li.p[p]];
        y[k] = s;
    }
    for (uint32_t r = 0; r < n; r++) x[r] = y[lu->pinv[r]];
}

uint32_t or_lu_n(const or_lu *lu) { return lu->n; }
uint32_t or_lu_nnz_l(const or_lu *lu) { return (uint32_t)lu->li.len; }
uint32_t or_lu_nnz_u(const or_lu *lu) { return (uint32_t)lu->ui.len; }
const uint32_t *or_lu_pinv(const or_lu *lu) { return lu->pinv; }
const uint32_t *or_lu_lp(const or_lu *lu) { return lu->lp; }
const uint32_t *or_lu_li(const or_lu *lu) { return lu->li.p; }
const double *or_lu_lx(const or_lu *lu) { return lu->lx.p; }
const uint32_t *or_lu_up(const or_lu *lu) { return lu->up; }
const uint32_t *or_lu_ui(const or_lu *lu) { return lu->ui.p; }
const double *or_lu_ux(const or_lu *lu) { return lu->ux.p; }
const double *or_lu_udiag(const or_lu *lu) { return lu->udiag; }
const uint32_t *or_lu_prow(const or_lu *lu) { return lu->prow; }
const uint8_t *or_lu_void_col(const or_lu *lu) { return lu->void_col; }
const uint8_t *or_lu_scaled_pivot(const or_lu *lu) { return lu->scaled_pivot; }
uint32_t or_lu_n_void_slots(const or_lu *lu) { return (uint32_t)lu->void_slots.len; }
const uint32_t *or_lu_void_slots(const or_lu *lu) { return lu->void_slots.p; }
const double *or_lu_w(const or_lu *lu) { return lu->w; }
int or_lu_sn_reached(const or_lu *lu) { return lu->sn_reached; }
