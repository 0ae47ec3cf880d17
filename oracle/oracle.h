/* CPU oracle of the ESPice GPU lane path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP lane engine
 * (egspice_amd/) is compared against. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it. The product path never links,
 * loads or calls anything here.
 *
 * It restates the reference CPU path in plain C, one lane at a time:
 *   omath.c      exp/log/pow bodies shared by every device evaluation
 *   order.c      src/solver/order.zig  (BTF by Tarjan, AMD, cholCount)
 *   sparse_lu.c  src/solver/sparse_lu.zig (Gilbert-Peierls factor, replay, solve)
 *   direct.c     src/solver/direct.zig (engine choice, bypass, replay probe)
 *   lane.c       src/solver/kernels/lane.zig (LaneTab, per-lane bodies, dense peel)
 *   devices.c    models/{vsource,resistor,capacitor,diode,mos1}.va bodies
 *   circuit.c    src/device/Circuit.zig freeze + src/analysis/Circuit.zig planes,
 *                src/solver/converger.zig newton/finalizeStep and the lane
 *                Newton of src/solver/lane_device.zig, the DC sweep
 *   tran.c       src/analysis/tran/tran.zig Sim (step control, companion,
 *                LTE, PULSE breakpoints) and the lane step machine
 *   meas.c       src/output/scan.zig .meas passes
 *   mc.c         the Monte Carlo sampler (netlist.siteKey, Sampler.uniform,
 *                normalQuantile, TrialDraw.sample)
 *
 * Parity status of each function is recorded in the header comment of its
 * file ("parity pinned by ..." or "parity unpinned").
 */
#ifndef ESPICE_ORACLE_H
#define ESPICE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NONE 0xffffffffu

/* ---------------- omath.c ---------------- */
double om_exp(double x);
double om_log(double x);
double om_pow(double x, double y);
double om_sin(double x);
double om_min(double a, double b);
double om_max(double a, double b);

/* ---------------- order.c ---------------- */
/* q[k] = column factored at step k; 0 ok, -1 out of workspace. */
int or_order(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *q);
/* AMD alone on the whole pattern (order.zig `amd`). */
int or_amd(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *q);
/* Tarjan blocks, sink first; returns block count, -1 on workspace. */
int or_blocks(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *emit, uint32_t *block_ptr);
/* Cholesky count of A + A^T under perm, stopping at cap. */
uint64_t or_chol_count(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const uint32_t *perm, uint64_t cap);
uint64_t or_diag_fill(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const uint32_t *q);

/* ---------------- sparse_lu.c ---------------- */
typedef struct or_lu or_lu;
or_lu *or_lu_new(uint32_t n, const uint32_t *col_ptr, const uint32_t *q);
void or_lu_free(or_lu *lu);
/* 0 ok, 1 SingularMatrix. */
int or_lu_factor(or_lu *lu, const uint32_t *col_ptr, const uint32_t *row_idx, const double *vals, double pivot_tol);
int or_lu_refactor(or_lu *lu, const uint32_t *col_ptr, const double *vals, double growth_limit);
void or_lu_solve(const or_lu *lu, const double *b, double *x);
void or_lu_solve_t(const or_lu *lu, const double *b, double *x);
void or_lu_set_caps(or_lu *lu, double fill_cap, double keep_min, uint64_t diag_fill);
/* Read-only views of the factor (valid until the next factor). */
uint32_t or_lu_n(const or_lu *lu);
uint32_t or_lu_nnz_l(const or_lu *lu);
uint32_t or_lu_nnz_u(const or_lu *lu);
const uint32_t *or_lu_pinv(const or_lu *lu);
const uint32_t *or_lu_lp(const or_lu *lu);
const uint32_t *or_lu_li(const or_lu *lu);
const double *or_lu_lx(const or_lu *lu);
const uint32_t *or_lu_up(const or_lu *lu);
const uint32_t *or_lu_ui(const or_lu *lu);
const double *or_lu_ux(const or_lu *lu);
const double *or_lu_udiag(const or_lu *lu);
const uint32_t *or_lu_prow(const or_lu *lu);
const uint8_t *or_lu_void_col(const or_lu *lu);
const uint8_t *or_lu_scaled_pivot(const or_lu *lu);
uint32_t or_lu_n_void_slots(const or_lu *lu);
const uint32_t *or_lu_void_slots(const or_lu *lu);
const double *or_lu_w(const or_lu *lu);

/* ---------------- direct.c ---------------- */
#define OR_PIVOT_TOL 1e-3
#define OR_GROWTH_LIMIT 1e-12
int or_is_tridiag(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx);
typedef struct or_solver or_solver;
/* NULL when the pattern is tridiagonal (that engine has no lane tape and
 * is not restated) or ordering ran out of workspace. */
or_solver *or_solver_new(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx);
or_solver *or_solver_new_ordered(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const uint32_t *q,
                                 uint64_t diag_fill);
void or_solver_free(or_solver *s);
int or_solver_factor(or_solver *s, const double *vals);
void or_solver_solve_neg(or_solver *s, const double *rhs, double *dx);
or_lu *or_solver_lu(or_solver *s);
const uint32_t *or_solver_q(const or_solver *s);

/* ---------------- lane.c ---------------- */
typedef struct {
    uint32_t n, nnz, nnz_l, nnz_u;
    double growth, wide;
    uint32_t col_ptr, prow, q, pinv, lp, li, up, ui, flags, void_slots, n_void_slots, _pad;
} or_lane_tab;
/* Flattens a factored LU into idx (caller frees with free()). */
uint32_t *or_build_tab(const or_lu *lu, const uint32_t *col_ptr, const uint32_t *q, double growth, double pivot_tol, or_lane_tab *tab, size_t *len);
int or_lane_refactor(const or_lane_tab *t, const uint32_t *x, const double *a, double *lx, double *ux, double *udiag, double *w, uint32_t N, uint32_t l);
void or_lane_solve_neg(const or_lane_tab *t, const uint32_t *x, const double *lx, const double *ux, const double *udiag, const double *rhs, double *y, double *dx, uint32_t N, uint32_t l);
double or_lane_update_norm(uint32_t n, double *x, const double *dx, double *x_old, const uint8_t *cur_row, double reltol, double abstol, double vntol, uint32_t N, uint32_t l);
int or_lane_plausible(uint32_t n, const double *x, double limit, uint32_t N, uint32_t l);
int or_lane_residual(uint32_t n, double reltol, double vntol, double residual_tol, const double *x, const double *rhs, const double *vals, const uint32_t *diag, uint32_t N, uint32_t l);
double or_lane_plausible_limit(uint32_t n, const double *x, const double *f, const double *vals, const uint32_t *diag, uint32_t N, uint32_t l);
int or_dense_solve_neg(uint32_t n, double *m, double *x);

/* ---------------- circuit / devices / newton ---------------- */
#include "oracle_circuit.h"

#ifdef __cplusplus
}
#endif
#endif
