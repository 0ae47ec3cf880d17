/* Circuit template, devices and Newton drivers of the oracle. TEST
 * INFRASTRUCTURE ONLY (see oracle.h). Included by oracle.h. */
#ifndef ESPICE_ORACLE_CIRCUIT_H
#define ESPICE_ORACLE_CIRCUIT_H

#include <stdint.h>

/* Device kinds, in the fixed order their parameter tables are declared.
 * Evaluation order is the order kinds first appear in the netlist
 * (src/device/Circuit.zig: one batch per proto, protos in first-use
 * order), instances in card order within a kind. */
enum { OR_VSRC = 0, OR_ISRC = 1, OR_RES = 2, OR_CAP = 3, OR_DIODE = 4, OR_MOS1 = 5, OR_PSP103 = 6, OR_BSIM4 = 7, OR_NKIND = 8 };
/* OR_PSP103 is compiled from models/psp103.va (va.c): its parameters are the
 * module's, in declaration order (or_va_param_index), NaN = not given. */
int or_va_param_index(int kind, const char *name);
int or_va_nparam(int kind);
int or_va_nu(int kind);

/* User parameter slots per kind (NaN = not given, the default applies). */
enum { OR_VSRC_DC = 0, OR_VSRC_WAVE, OR_VSRC_V1, OR_VSRC_V2, OR_VSRC_TD, OR_VSRC_TR, OR_VSRC_TF, OR_VSRC_PW, OR_VSRC_PER,
       OR_VSRC_SVO, OR_VSRC_SVA, OR_VSRC_SFREQ, OR_VSRC_STD, OR_VSRC_STHETA, OR_VSRC_SPHASE, OR_VSRC_PWL_LEN,
       OR_VSRC_PWL_R, OR_VSRC_PWL_TD, OR_VSRC_PWL_T, OR_VSRC_PWL_V = OR_VSRC_PWL_T + 16,
       OR_VSRC_EXP_V1 = OR_VSRC_PWL_V + 16, OR_VSRC_EXP_V2, OR_VSRC_EXP_TD1, OR_VSRC_EXP_TAU1, OR_VSRC_EXP_TD2,
       OR_VSRC_EXP_TAU2, OR_VSRC_SFFM_VO, OR_VSRC_SFFM_VA, OR_VSRC_SFFM_FM, OR_VSRC_SFFM_MDI, OR_VSRC_SFFM_FC,
       OR_VSRC_SFFM_TD, OR_VSRC_SFFM_PHASEM, OR_VSRC_SFFM_PHASEC, OR_VSRC_AM_VA, OR_VSRC_AM_VO, OR_VSRC_AM_MF,
       OR_VSRC_AM_FC, OR_VSRC_AM_TD, OR_VSRC_AM_PHASEC, OR_VSRC_AM_PHASES, OR_VSRC_K };
enum { OR_ISRC_DC = 0, OR_ISRC_K = OR_VSRC_K }; /* the voltage source's slots (isource.va:63-130) */
enum { OR_RES_R = 0, OR_RES_K };
enum { OR_CAP_C = 0, OR_CAP_K };
enum { OR_DIO_IS = 0, OR_DIO_N, OR_DIO_CJO, OR_DIO_VJ, OR_DIO_M, OR_DIO_FC, OR_DIO_TT, OR_DIO_GMIN, OR_DIO_AREA,
       OR_DIO_TNOM, OR_DIO_TEMP, OR_DIO_DTEMP, OR_DIO_EG, OR_DIO_XTI, OR_DIO_RS, OR_DIO_TRS, OR_DIO_TRS2, OR_DIO_TM1,
       OR_DIO_TM2, OR_DIO_K };
enum { OR_M1_TYPE = 0, OR_M1_VTO, OR_M1_KP, OR_M1_GAMMA, OR_M1_PHI, OR_M1_LAMBDA, OR_M1_TOX, OR_M1_LD, OR_M1_U0,
       OR_M1_NSUB, OR_M1_NSS, OR_M1_TPG, OR_M1_IS, OR_M1_JS, OR_M1_CBD, OR_M1_CBS, OR_M1_CJ, OR_M1_CJSW, OR_M1_PB,
       OR_M1_MJ, OR_M1_MJSW, OR_M1_FC, OR_M1_CGSO, OR_M1_CGDO, OR_M1_CGBO, OR_M1_TNOM, OR_M1_GMIN, OR_M1_TEMP,
       OR_M1_DTEMP, OR_M1_L, OR_M1_W, OR_M1_AD, OR_M1_AS, OR_M1_PD, OR_M1_PS, OR_M1_M, OR_M1_K };

typedef struct or_circuit or_circuit;
typedef struct or_lanes or_lanes;

or_circuit *or_ckt_new(void);
void or_ckt_free(or_circuit *c);
/* Row of a named net, allocated on first touch ("0" and "gnd" are row 0),
 * as frontend/builder.zig:559-565 rowOf does. */
uint32_t or_ckt_node(or_circuit *c, const char *name);
/* Adds one instance; `ports` are rows from or_ckt_node; internal unknowns
 * (the vsource branch current) get fresh rows now, builder.zig:238-266.
 * Returns the instance index within its kind, or -1. */
int32_t or_ckt_add(or_circuit *c, int kind, const uint32_t *ports, const double *params);
/* Builds the CSC pattern (every diagonal, then each device's structural
 * entries, device/Circuit.zig:112-201), diag slots and slot tapes. */
int or_ckt_freeze(or_circuit *c);
uint32_t or_ckt_n(const or_circuit *c);
uint32_t or_ckt_nnz(const or_circuit *c);
const uint32_t *or_ckt_col_ptr(const or_circuit *c);
const uint32_t *or_ckt_row_idx(const or_circuit *c);
const uint32_t *or_ckt_diag(const or_circuit *c);
const uint8_t *or_ckt_cur_row(const or_circuit *c);
uint32_t or_ckt_count(const or_circuit *c, int kind);

or_lanes *or_lanes_new(const or_circuit *c, uint32_t n_lanes);
void or_lanes_free(or_lanes *L);
/* Per-lane user parameter values (n_lanes entries). */
void or_lanes_set_param(or_lanes *L, int kind, uint32_t inst, uint32_t k, const double *values);
/* Temperature/geometry preparation of every instance on every lane. */
void or_lanes_prepare(or_lanes *L);

/* Newton options (converger.Options subset). */
typedef struct {
    double reltol, abstol, vntol, residual_tol, gmin;
    uint32_t max_iter;
    uint32_t init_fix;
} or_newton_opts;

/* Lane outcome: bit 0 converged, bit 1 singular; iterations as
 * converger.Result.iterations. */
typedef struct { uint32_t status; uint32_t iterations; } or_outcome;

/* Static (OP) Newton on every lane from x (lane-major n*N, in/out).
 * mode 0: the GPU lane algorithm (lane 0's tape, lane bodies, dense peel),
 *         src/analysis/gpu/lanes.zig newton/ticks + solver/lane_device.zig.
 * mode 1: the reference CPU path per lane, converger.newton with its own
 *         direct.Solver (full factor, refactor, replay probe).
 * Returns 0, or -1 when the template has no lane tape (tridiagonal
 * pattern or a singular lead factor). */
int or_newton_op(or_lanes *L, int mode, const or_newton_opts *o, double *x, or_outcome *out);

/* Transient of one lane (tran.c): `.tran dt_init t_stop` with the step
 * ceiling dt_max already resolved (the deck rule is min(dt_init,
 * t_stop/50)); method 0 BE, 1 trapezoidal, 2 gear-2. */
typedef struct {
    double t_stop, dt_init, dt_max, dt_min, xmu;
    double reltol, abstol, vntol, chgtol, trtol, residual_tol_op;
    uint32_t method, max_iter, max_steps, _pad;
} or_tran_opts;

typedef struct {
    uint32_t points, steps, attempts, nr_iters, rej_newton, rej_lte, bp_landings;
    int32_t status; /* 0 reached t_stop, 1 dt underflow, 2 no operating point, 3 out of memory */
} or_tran_stats;

/* Records every accepted row: times[k] and rows[k*n + i] (malloc'd, the
 * caller frees them). Returns the row count. */
void or_tran_coeffs(int method, double dt, double dt_prev, double xmu, double *out);
int or_tran_run(or_lanes *L, uint32_t l, const or_tran_opts *o, double **times, double **rows, or_tran_stats *st);
/* The same from a warm operating point at x0 (n values); x0 NULL is cold. */
int or_tran_run_from(or_lanes *L, uint32_t l, const or_tran_opts *o, const double *x0, double **times, double **rows,
                     or_tran_stats *st);

/* .meas passes over a recorded waveform (meas.c). A clause's counts use
 * -1 for not given and -2 for LAST. */
typedef struct {
    double from, to, val, td;
    int32_t rise, fall, cross, _pad;
} or_meas_clause;

/* The event a WHEN/TRIG/TARG clause names; NaN when it never happens. */
double or_meas_cross(const double *x, const double *y, uint32_t len, const or_meas_clause *c);
/* y at `at` (slope 0) or the slope of the segment holding it (slope 1). */
double or_meas_at(const double *x, const double *y, uint32_t len, double at, int slope);
double or_meas_find_when(const double *x, const double *y, const double *y2, uint32_t len, const or_meas_clause *cl,
                         int slope);
/* op 0 MIN, 1 MAX (value, where), 2 AVG (value, last scale read). */
void or_meas_extremum(const double *x, const double *y, uint32_t len, double from, double to, int op, double *value,
                      double *at);
/* INTEG (rms 0) or RMS (rms 1) over [from, to]. */
void or_meas_integ(const double *x, const double *y, uint32_t len, double from, double to, int rms, double *value,
                   double *rfrom, double *rto);

/* Monte Carlo draws (mc.c). */
uint64_t or_site_key(uint64_t owner, uint64_t salt);
double or_mc_uniform(uint64_t seed, uint64_t site, uint32_t trial);
double or_normal_quantile(double u);
double or_trial_draw(uint64_t seed, uint32_t trial, uint64_t site, int fn, const double *args, uint32_t nargs);
uint64_t or_wyhash(uint64_t seed, const char *key, size_t len);
uint64_t or_param_site(const char *name, uint32_t ordinal);
/* The same on `threads` POSIX threads, lanes dealt round robin. */
int or_newton_op_mt(or_lanes *L, int mode, const or_newton_opts *o, double *x, or_outcome *out, uint32_t threads);
/* The lead tape of lane 0 at x (mode 0's tape), for tests. */
int or_lead_tab(or_lanes *L, const double *x, int init_fix, or_lane_tab *tab, uint32_t **idx, size_t *len);
/* One residual+Jacobian evaluation of lane l at x (n entries), unlimited,
 * ground pin included: rhs (n), vals (nnz). For the FD Jacobian tests. */
void or_eval_lane(or_lanes *L, uint32_t l, const double *x, double *rhs, double *vals);

/* `.dc` sweep of one vsource's dc value on lane 0 (src/analysis/dc/dc.zig
 * run/Sweep: cold first point, predictor 2 x_k - x_{k-1} limited against
 * the last point, warm Newton per point). x_out: npts*n, point-major. */
int or_dc_sweep(or_lanes *L, uint32_t vsrc, double start, double step, uint32_t npts, const or_newton_opts *o, double *x_out, uint32_t *converged);

/* The operating point's ladder after the plain rung (ladder.c,
 * src/analysis/dc/op.zig:140-415). */
/* One lane's static Newton under a homotopy: gmin o->gmin at `stamp`
 * (2n words; NULL: every diagonal), sources times srcfact, cold when
 * o->init_fix; the lane's ordered direct solver. */
int or_newton_hom(or_lanes *L, uint32_t l, const or_newton_opts *o, double *xl, or_outcome *out, const uint32_t *stamp,
                  double srcfact);
int or_newton_hom_lim(or_lanes *L, uint32_t l, const or_newton_opts *o, double *xl, or_outcome *out,
                      const uint32_t *stamp, double srcfact, int lim_active);
/* A `.dc` sweep of source (kind, inst)'s dc value over every lane, one lane
 * at a time (dc.zig:161-530): warm from the predictor limited against the
 * last point, else the cold ladder. x_out[npts][n][N] (NaN: no
 * convergence), converged[npts][N]. */
int or_dc_sweep_lanes(or_lanes *L, int kind, uint32_t inst, const double *values, uint32_t npts,
                      const or_newton_opts *plain, double gmin_start, uint32_t budget, double *x_out,
                      uint32_t *converged);
/* Lane l's G and C planes at x (nnz each; the AC linearization). */
void or_eval_lane_gc(or_lanes *L, uint32_t l, const double *x, double *g_out, double *c_out);
/* Lane l's compiled instances set up at $simparam("gmin") = gmin. */
void or_lanes_set_model_gmin(or_lanes *L, uint32_t l, double gmin);
/* gminStamps from lane l's planes at x = 0 (2n words). */
void or_gmin_stamps(or_lanes *L, uint32_t l, uint32_t *stamp);
double or_gmin_factor(double factor, uint32_t iterations, uint32_t budget);
int or_gmin_stalled(const double *gmin, const int *conv, uint32_t count, double start, double next);
uint32_t or_gmin_replay(const int *conv, const uint32_t *it, uint32_t count, double gmin_start, uint32_t budget,
                        double *next, int *action);
/* Plain rung then rungs 2-3 and the final Newton per failing lane; x
 * (n * N node-major) in and out; out[l] status 1 / 0, iterations summed
 * (saturating at 65535); rung[l] 1 plain, 2 gmin, 3 source, 0 none. */
int or_op_ladder(or_lanes *L, const or_newton_opts *plain, double gmin_start, uint32_t budget, double *x,
                 or_outcome *out, uint32_t *rung);

#endif
