/* Per-lane bodies oracle: restates src/solver/kernels/lane.zig. TEST
 * INFRASTRUCTURE ONLY (see oracle.h). Lane l of N holds element i at
 * i*N + l (lane-major), exactly as the device planes do.
 *
 *   or_build_tab             lane.zig:778-802 (LaneTab words: col_ptr,
 *                            prow, q, pinv, lp, li, up, ui, void_slots,
 *                            then n flag words: bit 0 void, bit 1 scaled)
 *   or_lane_refactor         lane.zig:56-125 (refactor + wideMultipliers)
 *   or_lane_solve_neg        lane.zig:130-157
 *   or_lane_update_norm      lane.zig:161-179 (converger.updateAndNorm W=1)
 *   or_lane_plausible        lane.zig:182-185
 *   or_lane_residual         lane.zig:189-198 (converger.residualGate)
 *   or_lane_plausible_limit  lane.zig:202-214 (converger.plausibleLimit)
 *   or_dense_solve_neg       lane.zig:310-350 (dense partial-pivot LU)
 * Parity pinned by: tests/test_oracle_lane.py, which replays
 * lane.zig:811-908 (bodies bitwise the scalar SparseLu refactor and
 * solve, failing exactly where the replay fails or a multiplier runs wide)
 * against this file's own sparse_lu.c; the dense solve against
 * numpy.linalg.solve (tests/test_oracle_lu.py). The update, plausible,
 * residual and plausible-limit forms are the Newton driver's own
 * (circuit.c), pinned with it by tests/test_oracle_newton.py (the
 * diode_iv golden rows, KCL at every converged lane).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint64_t ubits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }
static double from_ubits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }

uint32_t *or_build_tab(const or_lu *lu, const uint32_t *col_ptr, const uint32_t *q, double growth, double pivot_tol, or_lane_tab *t, size_t *len_out)
{
    const uint32_t n = or_lu_n(lu);
    const uint32_t nnz = col_ptr[n];
    const uint32_t nl = or_lu_nnz_l(lu), nu = or_lu_nnz_u(lu), nv = or_lu_n_void_slots(lu);
    memset(t, 0, sizeof(*t));
    t->n = n;
    t->nnz = nnz;
    t->nnz_l = nl;
    t->nnz_u = nu;
    t->growth = growth;
    t->wide = 1 / pivot_tol;
    t->n_void_slots = nv;
    size_t len = (size_t)(n + 1) + nnz + n + n + (n + 1) + nl + (n + 1) + nu + nv + n;
    uint32_t *idx = malloc(sizeof(uint32_t) * (len ? len : 1));
    uint32_t at = 0;
#define PART(field, src, cnt) do { t->field = at; if (cnt) memcpy(idx + at, src, sizeof(uint32_t) * (cnt)); at += (cnt); } while (0)
    PART(col_ptr, col_ptr, n + 1);
    PART(prow, or_lu_prow(lu), nnz);
    PART(q, q, n);
    PART(pinv, or_lu_pinv(lu), n);
    PART(lp, or_lu_lp(lu), n + 1);
    PART(li, or_lu_li(lu), nl);
    PART(up, or_lu_up(lu), n + 1);
    PART(ui, or_lu_ui(lu), nu);
    PART(void_slots, or_lu_void_slots(lu), nv);
#undef PART
    t->flags = at;
    const uint8_t *vc = or_lu_void_col(lu), *sp = or_lu_scaled_pivot(lu);
    for (uint32_t k = 0; k < n; k++) idx[at + k] = (vc[k] ? 1u : 0u) | (sp[k] ? 2u : 0u);
    *len_out = len;
    return idx;
}

int or_lane_refactor(const or_lane_tab *t, const uint32_t *x, const double *a, double *lx, double *ux, double *udiag, double *w, uint32_t N_, uint32_t l)
{
    const size_t N = N_, L = l;
    int bad = 0;
    uint64_t nonfinite = 0;
    for (uint32_t v = 0; v < t->n_void_slots; v++)
        if (a[(size_t)x[t->void_slots + v] * N + L] != 0) return 1;
    for (uint32_t k = 0; k < t->n; k++) {
        uint32_t c = x[t->q + k];
        for (uint32_t p = x[t->col_ptr + c]; p < x[t->col_ptr + c + 1]; p++) w[(size_t)x[t->prow + p] * N + L] = a[(size_t)p * N + L];
        for (uint32_t p = x[t->up + k]; p < x[t->up + k + 1]; p++) {
            size_t i = x[t->ui + p];
            double uki = w[i * N + L];
            w[i * N + L] = 0;
            ux[(size_t)p * N + L] = uki;
            nonfinite |= ubits(uki - uki);
            for (uint32_t pl = x[t->lp + i]; pl < x[t->lp + i + 1]; pl++)
                w[(size_t)x[t->li + pl] * N + L] -= lx[(size_t)pl * N + L] * uki;
        }
        if (x[t->flags + k] & 1u) {
            udiag[(size_t)k * N + L] = 1;
            for (uint32_t p = x[t->lp + k]; p < x[t->lp + k + 1]; p++) {
                lx[(size_t)p * N + L] = 0;
                w[(size_t)x[t->li + p] * N + L] = 0;
            }
            w[(size_t)k * N + L] = 0;
            for (uint32_t p = x[t->up + k]; p < x[t->up + k + 1]; p++) {
                uint32_t i = x[t->ui + p];
                for (uint32_t pl = x[t->lp + i]; pl < x[t->lp + i + 1]; pl++) w[(size_t)x[t->li + pl] * N + L] = 0;
            }
            continue;
        }
        double d = w[(size_t)k * N + L];
        w[(size_t)k * N + L] = 0;
        int dead = d == 0 || !(fabs(d) < INFINITY);
        bad = bad || dead;
        if (dead) d = 1;
        udiag[(size_t)k * N + L] = d;
        if (t->growth > 0 && (x[t->flags + k] & 2u) == 0) {
            int64_t cbits = (int64_t)ubits(fabs(d));
            for (uint32_t p = x[t->lp + k]; p < x[t->lp + k + 1]; p++) {
                size_t r = x[t->li + p];
                double v = w[r * N + L];
                w[r * N + L] = 0;
                int64_t b = (int64_t)ubits(fabs(v));
                if (b > cbits) cbits = b;
                lx[(size_t)p * N + L] = v / d;
            }
            double cmax = from_ubits((uint64_t)cbits);
            bad = bad || !(fabs(d) >= t->growth * cmax);
        } else {
            for (uint32_t p = x[t->lp + k]; p < x[t->lp + k + 1]; p++) {
                size_t r = x[t->li + p];
                double lv = w[r * N + L] / d;
                nonfinite |= ubits(lv - lv);
                lx[(size_t)p * N + L] = lv;
                w[r * N + L] = 0;
            }
        }
    }
    if (nonfinite != 0) return 1;
    uint64_t worst = 0;
    for (uint32_t k = 0; k < t->n; k++) {
        if (x[t->flags + k] & 2u) continue;
        for (uint32_t p = x[t->lp + k]; p < x[t->lp + k + 1]; p++) {
            uint64_t b = ubits(fabs(lx[(size_t)p * N + L]));
            if (b > worst) worst = b;
        }
    }
    return bad || !(from_ubits(worst) <= t->wide);
}

void or_lane_solve_neg(const or_lane_tab *t, const uint32_t *x, const double *lx, const double *ux, const double *udiag, const double *rhs, double *y, double *dx, uint32_t N_, uint32_t l)
{
    const size_t N = N_, L = l;
    for (uint32_t r = 0; r < t->n; r++) y[(size_t)x[t->pinv + r] * N + L] = -rhs[(size_t)r * N + L];
    for (uint32_t k = 0; k < t->n; k++) {
        double yk = y[(size_t)k * N + L];
        if (yk != 0)
            for (uint32_t p = x[t->lp + k]; p < x[t->lp + k + 1]; p++) y[(size_t)x[t->li + p] * N + L] -= lx[(size_t)p * N + L] * yk;
    }
    for (uint32_t k = t->n; k > 0;) {
        k--;
        double zk = y[(size_t)k * N + L] / udiag[(size_t)k * N + L];
        y[(size_t)k * N + L] = zk;
        if (zk != 0)
            for (uint32_t p = x[t->up + k]; p < x[t->up + k + 1]; p++) y[(size_t)x[t->ui + p] * N + L] -= ux[(size_t)p * N + L] * zk;
    }
    for (uint32_t j = 0; j < t->n; j++) dx[(size_t)x[t->q + j] * N + L] = y[(size_t)j * N + L];
}

double or_lane_update_norm(uint32_t n, double *x, const double *dx, double *x_old, const uint8_t *cur_row, double reltol, double abstol, double vntol, uint32_t N_, uint32_t l)
{
    const size_t N = N_, L = l;
    int64_t worst = 0;
    for (uint32_t i = 0; i < n; i++) {
        double xo = x[i * N + L];
        double d = dx[i * N + L];
        double xn = xo + d;
        x_old[i * N + L] = xo;
        x[i * N + L] = xn;
        /* Zig's float @max is LLVM maxnum: a NaN operand yields the other */
        double tol = reltol * fmax(fabs(xn), fabs(xo)) + (cur_row[i] ? abstol : vntol);
        int finite = fabs(xn) < INFINITY && fabs(d) < INFINITY;
        double q = fabs(d) / tol;
        double score = !finite ? INFINITY : (q == q ? q : 0);
        int64_t b = (int64_t)ubits(score);
        if (b > worst) worst = b;
    }
    return from_ubits((uint64_t)worst);
}

int or_lane_plausible(uint32_t n, const double *x, double limit, uint32_t N, uint32_t l)
{
    for (uint32_t i = 0; i < n; i++)
        if (!(fabs(x[(size_t)i * N + l]) <= limit)) return 0;
    return 1;
}

int or_lane_residual(uint32_t n, double reltol, double vntol, double residual_tol, const double *x, const double *rhs, const double *vals, const uint32_t *diag, uint32_t N, uint32_t l)
{
    for (uint32_t i = 0; i < n; i++) {
        double scale = fabs(vals[(size_t)diag[i] * N + l]);
        double s = scale == 0 ? 1 : scale;
        double tol = fmax(residual_tol, 10.0 * s * (reltol * fabs(x[(size_t)i * N + l]) + vntol));
        if (!(fabs(rhs[(size_t)i * N + l]) <= tol)) return 0;
    }
    return 1;
}

double or_lane_plausible_limit(uint32_t n, const double *x, const double *f, const double *vals, const uint32_t *diag, uint32_t N, uint32_t l)
{
    double g = 1;
    for (uint32_t i = 0; i < n; i++) g = fmax(g, fabs(vals[(size_t)diag[i] * N + l]));
    uint64_t xm = 0, fm = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint64_t bx = ubits(fabs(x[(size_t)i * N + l]));
        uint64_t bf = ubits(fabs(f[(size_t)i * N + l]));
        if (bx > xm) xm = bx;
        if (bf > fm) fm = bf;
    }
    double scale = fmin(fmax(from_ubits(xm), from_ubits(fm) * g), 1e18);
    return !(1e3 * scale > 1e9) ? 1e9 : 1e3 * scale;
}

int or_dense_solve_neg(uint32_t n, double *m, double *x)
{
    const double tiny = 2.220446049250313e-16 * 2.220446049250313e-16;
    for (uint32_t i = 0; i < n; i++) x[i] = -x[i];
    for (uint32_t k = 0; k < n; k++) {
        uint32_t max_row = k;
        double max_val = fabs(m[(size_t)k * n + k]);
        for (uint32_t i = k + 1; i < n; i++) {
            double v = fabs(m[(size_t)i * n + k]);
            if (v > max_val) {
                max_val = v;
                max_row = i;
            }
        }
        double p = m[(size_t)max_row * n + k];
        if (!(fabs(p) >= tiny && fabs(p) <= 1.7976931348623157e308)) return 0;
        if (max_row != k) {
            for (uint32_t j = 0; j < n; j++) {
                double t = m[(size_t)k * n + j];
                m[(size_t)k * n + j] = m[(size_t)max_row * n + j];
                m[(size_t)max_row * n + j] = t;
            }
            double t = x[k];
            x[k] = x[max_row];
            x[max_row] = t;
        }
        double inv = 1.0 / m[(size_t)k * n + k];
        for (uint32_t i = k + 1; i < n; i++) {
            double f = m[(size_t)i * n + k] * inv;
            for (uint32_t j = k + 1; j < n; j++) m[(size_t)i * n + j] -= f * m[(size_t)k * n + j];
            x[i] -= f * x[k];
        }
    }
    for (uint32_t ki = n; ki > 0;) {
        ki--;
        double s = 0;
        for (uint32_t j = ki + 1; j < n; j++) s += m[(size_t)ki * n + j] * x[j];
        x[ki] = (x[ki] - s) / m[(size_t)ki * n + ki];
    }
    return 1;
}
