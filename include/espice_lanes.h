/* espice lane engine C ABI: many instances (lanes) of one circuit template
 * solved together on an MI355X (gfx950), one 64-wide workgroup per lane
 * with the lane's whole state in LDS.
 *
 * This is the seam the reference's lane driver calls in place of its Zig
 * GPU lane backend: `GpuLanes` (src/analysis/gpu/lanes.zig) and its
 * exported `arp_lane_*` kernels (src/solver/lane_device.zig:623-668).
 * Plain pointers and sizes only; every call returns an espice_status
 * (include/espice.h's codes); a lane's numerical failure is a per-lane
 * result, never an error. Calls on one handle must be serialized; distinct
 * handles are independent. No call keeps a caller pointer after it
 * returns. The engine owns its device memory and one HIP stream per lane
 * batch.
 *
 * Layout: lane-major, element i of lane l of N at i * N + l, as the
 * reference's lane planes (lane_device.zig:1-6).
 */
#ifndef ESPICE_LANES_H
#define ESPICE_LANES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef ESPICE_H
typedef uint32_t espice_status;
enum {
    ESPICE_OK = 0, ESPICE_INVALID_ARGUMENT = 1, ESPICE_BUFFER_TOO_SMALL = 2,
    ESPICE_OUT_OF_MEMORY = 3, ESPICE_INVALID_QUERY = 4,
    ESPICE_RESULT_UNAVAILABLE = 5, ESPICE_FAILED = 6, ESPICE_ABI_MISMATCH = 7,
    ESPICE_INVALID_HANDLE = 8
};
#endif

/* Device kinds of a template (models/vsource.va, isource.va, resistor.va,
 * capacitor.va, diode.va, mos1.va; psp103.va and bsim4va.va compiled from
 * their Verilog-A: ports D G S B, the module's parameters in declaration
 * order, named through espice_lanes_param_index). */
enum {
    ESPICE_LANES_VSOURCE = 0, ESPICE_LANES_ISOURCE = 1, ESPICE_LANES_RESISTOR = 2,
    ESPICE_LANES_CAPACITOR = 3, ESPICE_LANES_DIODE = 4, ESPICE_LANES_MOS1 = 5, ESPICE_LANES_PSP103 = 6,
    ESPICE_LANES_BSIM4 = 7
};
/* User parameters per kind (NaN = not given). VSOURCE: dc, waveform (0 DC,
 * 1 PULSE, 2 SIN, 3 EXP, 4 PWL, 5 SFFM, 6 AM), PULSE v1 v2 td tr tf pw per, SIN vo va freq td
 * theta phase, PWL point count, r=, td=, 16 times, 16 values (its table
 * holds 64), EXP v1 v2 td1 tau1 td2 tau2, SFFM vo va fm mdi fc td phasem
 * phasec, AM va vo mf fc td phasec phases (models/vsource.va:58-130); the
 * current source has the same slots (models/isource.va:63-130), SFFM excepted; a dc not given is the waveform's t = 0
 * value (frontend/builder.zig:1599-1605, 2934-2966). DIODE: is n cjo vj m fc
 * tt gmin area tnom temp dtemp eg xti rs trs trs2 tm1 tm2 (models/diode.va:
 * 40-115); with rs > 0 the diode takes a fresh internal row at add, and every
 * lane must then keep rs > 0 (and rs = 0 lanes stay at 0). */
enum {
    ESPICE_LANES_VSOURCE_PARAMS = 71, ESPICE_LANES_ISOURCE_PARAMS = 71, ESPICE_LANES_RESISTOR_PARAMS = 1,
    ESPICE_LANES_CAPACITOR_PARAMS = 1, ESPICE_LANES_DIODE_PARAMS = 19, ESPICE_LANES_MOS1_PARAMS = 36
};
/* Diode parameter slots that decide its topology. */
enum { ESPICE_LANES_DIODE_RS = 14 };

typedef struct espice_lanes_template espice_lanes_template;
typedef struct espice_lanes espice_lanes;

/* converger.Options as the lane Newton reads it
 * (src/solver/converger.zig:100-122); defaults reltol 1e-3, abstol 1e-12,
 * vntol 1e-6, residual_tol 1e-9, gmin 0, max_iter 100
 * (src/core/numerics.zig:378-399). init_fix 1: a cold operating point
 * (MODEINITJCT seeds, then one INITFIX pass). */
typedef struct {
    double reltol, abstol, vntol, residual_tol, gmin;
    uint32_t max_iter, init_fix;
} espice_newton_opts;

/* Per-lane outcome, as converger.LaneResult: status bit 0 converged,
 * bit 1 singular (the lane's factor failed); iterations as
 * converger.Result.iterations. */
typedef struct {
    uint32_t status, iterations;
} espice_lane_result;

/* `LaneTab` (src/solver/kernels/lane.zig:17-37): word offsets of a frozen
 * LU tape inside one index buffer, plus its scalars. */
typedef struct {
    uint32_t n, nnz, nnz_l, nnz_u;
    double growth, wide;
    uint32_t col_ptr, prow, q, pinv, lp, li, up, ui, flags, void_slots, n_void_slots, pad;
} espice_lane_tab;

/* The thread's last error name (NUL-terminated); empty when none. */
espice_status espice_lanes_error_message(char *buffer, size_t capacity, size_t *required);

/* ---- template: a circuit frozen into one MNA pattern -------------------
 * Replaces the frontend's builder for lane batches (frontend/builder.zig
 * rowOf :559-565, addDevice :238-266; device/Circuit.zig freeze :112-201). */
espice_status espice_lanes_template_create(espice_lanes_template **out);
espice_status espice_lanes_template_destroy(espice_lanes_template *t);
/* Row of a named net, allocated on first touch; "0" and "gnd" are row 0. */
espice_status espice_lanes_template_node(espice_lanes_template *t, const char *name, uint32_t *row);
/* User parameters of a kind (the count espice_lanes_template_add takes),
 * and the slot of a named parameter of a compiled kind (case-insensitive,
 * aliasparam names resolved; UnknownParam otherwise). Replace the
 * builder's binding of card names onto a model (device/bind.zig). */
espice_status espice_lanes_param_count(uint32_t kind, uint32_t *count);
espice_status espice_lanes_param_index(uint32_t kind, const char *name, uint32_t *index);
/* One instance of kind on `ports` (rows); params holds the kind's user
 * parameters (NaN = not given). Internal unknowns take fresh rows now. */
espice_status espice_lanes_template_add(espice_lanes_template *t, uint32_t kind, const uint32_t *ports,
    const double *params, size_t param_count, uint32_t *instance);
/* Freezes the pattern: diagonals plus each device's structural entries,
 * the diagonal slots, the BTF+AMD column order (src/solver/order.zig). */
espice_status espice_lanes_template_freeze(espice_lanes_template *t);
espice_status espice_lanes_template_info(const espice_lanes_template *t, uint32_t *n, uint32_t *nnz,
    uint32_t *tridiagonal);
/* Copies the frozen pattern; any pointer may be NULL. */
espice_status espice_lanes_template_pattern(const espice_lanes_template *t, uint32_t *col_ptr, uint32_t *row_idx,
    uint32_t *diag, uint8_t *current_row, uint32_t *q);
/* Where the operating point's stepping rungs load their diagonal gmin
 * (src/analysis/dc/op.zig:711-790 gminStamps: ngspice's spMNA_Preorder
 * swaps each voltage-source branch column with a symmetric +-1 pair, then
 * LoadGmin adds gmin at each row's (row, col[row])). col[n], slot[n]:
 * row i's column and pattern slot, or 0xffffffff where row i takes none
 * (row 0, ground, never does). */
espice_status espice_lanes_template_gmin_stamps(const espice_lanes_template *t, uint32_t *col, uint32_t *slot);

/* ---- a batch of lanes on the GPU (GpuLanes.initStatic, lanes.zig:1080) - */
espice_status espice_lanes_create(const espice_lanes_template *t, uint32_t n_lanes, espice_lanes **out);
espice_status espice_lanes_destroy(espice_lanes *l);
/* Returns every lane's solver to its fresh state, as direct.Solver.reset
 * (src/solver/direct.zig:183-207): the next solve factors from scratch. */
espice_status espice_lanes_reset(espice_lanes *l);
/* values[n_lanes]: user parameter `param` of instance `instance` per lane. */
espice_status espice_lanes_set_param(espice_lanes *l, uint32_t kind, uint32_t instance, uint32_t param,
    const double *values);
/* Temperature and geometry preparation of every instance on every lane,
 * then the upload. Required after set_param and before a solve. */
espice_status espice_lanes_prepare(espice_lanes *l);
/* A voltage or current source's dc value on every lane of a prepared
 * batch at once, no prepare needed: a `.dc` sweep's per-point install
 * (ParamRef.set and recomputeType of the swept source,
 * src/analysis/dc/dc.zig:442-457). */
espice_status espice_lanes_set_source(espice_lanes *l, uint32_t kind, uint32_t instance, double value);
/* x[n * n_lanes] lane-major: each lane's start point. */
espice_status espice_lanes_upload_x(espice_lanes *l, const double *x);
espice_status espice_lanes_download_x(espice_lanes *l, double *x);
/* The static (operating-point) Newton of every lane on the GPU, one
 * kernel: the plain rung of op.solveGpuLanes (src/analysis/dc/op.zig:626).
 * out[n_lanes]; kernel_ms (optional) receives the kernel's GPU time. */
espice_status espice_lanes_newton_op(espice_lanes *l, const espice_newton_opts *opts, espice_lane_result *out,
    double *kernel_ms);
/* How the last espice_lanes_newton_op (or espice_lanes_newton_learned)
 * launch factored each lane's first matrix on the workgroup-per-lane path:
 * *predicted lanes took the template's predicted pivot sequence (checked
 * against the full factor's pivot rule at every step, so bitwise the same
 * factor), *serial lanes ran the serial full factor. No counterpart in the
 * reference, whose lanes always run the full factor first
 * (src/solver/sparse_lu.zig:274-311); diagnostics only. */
espice_status espice_lanes_first_factors(espice_lanes *l, uint32_t *predicted, uint32_t *serial);
/* The static Newton of every lane with each tick's linear step in mixed
 * precision (the lane path's FP32 numeric LU and triangular solves,
 * recovered by FP64 iterative refinement; the reference's
 * src/solver/sparse_lu.zig:973-1322 refinement against its residual bars):
 * the FP64 matrix rounded to FP32, factored densely with partial
 * pivoting after an exact power-of-two equilibration, solved in FP32, and
 * refined with FP64 residuals r = -rhs - A dx until the residual meets the
 * FP64 solve's backward-error bar max|r| <= max|dx| ||A||inf sqrt(n) 2^-53
 * (LAPACK dsgesv's test), at most max_refine FP32 solves per tick; a tick
 * whose refinement overflows or runs out of solves takes the FP64 factor
 * instead. Device evaluation, the residual,
 * the update and every gate stay FP64, so a converged lane's operating
 * point lies inside the same Newton tolerance as the FP64 path's.
 * refine[2 * n_lanes] (optional): per lane the FP32 solves taken and the
 * ticks handed back to FP64. Lanes up to 64 unknowns, templates of the
 * hand-written kinds or with PSP103 (MixedUnsupported otherwise). */
espice_status espice_lanes_newton_mixed(espice_lanes *l, const espice_newton_opts *opts, uint32_t max_refine,
    espice_lane_result *out, uint32_t *refine, double *kernel_ms);
/* Unknowns per lane and lanes of a batch. */
espice_status espice_lanes_shape(const espice_lanes *l, uint32_t *n, uint32_t *n_lanes);

/* One lane's homotopy for espice_lanes_newton_homotopy: the operating
 * point's stepping rungs (src/analysis/dc/op.zig:185-330) solve with a
 * diagonal gmin shunt gshunt (S) at the gmin stamps, the compiled kinds'
 * $simparam("gmin") at gmodel (the reference reruns their setup,
 * Circuit.zig:1394-1405 setHomotopy, eval.zig:2596-2612) and every
 * independent source's value times srcfact ($simparam("sourceScaleFactor"),
 * models/vsource.va:333, isource.va:332). flags: bit 0 the lane runs,
 * bit 1 it starts cold (the limiters' MODEINITJCT seeds, then INITFIX)
 * rather than warm. */
typedef struct {
    double gshunt, gmodel, srcfact;
    uint32_t flags, pad;
} espice_lane_homotopy;
enum { ESPICE_HOM_ACTIVE = 1, ESPICE_HOM_COLD = 2, ESPICE_HOM_PREDICT = 4 };
/* The static Newton of the lanes whose control is active, each under its
 * homotopy, from its x (espice_lanes_upload_x); the others are untouched
 * and their out entries read {0, 0}. opts gives the tolerances and
 * max_iter; its gmin and init_fix are the controls'. A warm lane with
 * ESPICE_HOM_PREDICT first limits its start against x_prev (n * n_lanes,
 * node-major like x; required when any lane predicts), as a `.dc` point's
 * predictor is limited against the previous point
 * (src/analysis/dc/dc.zig:478-483 prepareWarm). A compiled kind's
 * gmodel lands in the prepared constants that hold its $simparam("gmin")
 * (found by rerunning its setup at a second gmin: HomotopyGminDerived when
 * a constant derives from it otherwise) and stays there: a lane's last
 * homotopy solve leaves its model gmin. HomotopyLaneTooLarge when the
 * lanes do not run workgroup-per-lane. */
espice_status espice_lanes_newton_homotopy(espice_lanes *l, const espice_newton_opts *opts,
    const espice_lane_homotopy *lanes, const double *x_prev, espice_lane_result *out, double *kernel_ms);

/* The operating-point ladder's options: gmin_start (1e-2 S, the gmin
 * descent's first good gmin, src/core/numerics.zig:390) and the Newton
 * budget of each stepping solve (100: optionsFromTolerances' floor of 100
 * over itl2 = 50, converger.zig:86-88). */
typedef struct {
    double gmin_start;
    uint32_t max_iter, pad;
} espice_ladder_opts;
/* Every lane's operating point by the reference's ladder
 * (src/analysis/dc/op.zig:140-330 solveLadder/afterPlain, the host half of
 * solveOnDevice :660-688): the plain rung on every lane (opts as
 * espice_lanes_newton_op), then, for each lane it leaves unconverged or
 * singular, dynamic gmin stepping from gmin_start down to the deck's gmin
 * (1e-12 S) and a clean solve without the shunt, else source stepping
 * from zero (with the sources-off gmin rescue) and a final Newton. The
 * lanes that need a rung run together, one homotopy launch per solve.
 * out[l]: status 1 converged, 0 not; iterations summed over every solve,
 * saturating at 65535. rung[l] (optional): 1 plain, 2 gmin stepping, 3
 * source stepping, 0 none converged (the reference's JFNK and OPtran rungs
 * are not built). Each lane's x is left on the device. */
espice_status espice_lanes_op_ladder(espice_lanes *l, const espice_newton_opts *opts, const espice_ladder_opts *lo,
    espice_lane_result *out, uint32_t *rung, double *kernel_ms);
/* A `.dc` sweep of one source's dc value over every lane
 * (src/analysis/dc/dc.zig:161-293 runGpuLanes with its Sweep :303-530):
 * point by point the source takes values[pt] on every lane; each lane that
 * converged at the last point starts from the last two points extrapolated
 * (MODEINITPRED; the first point's solution also stands for the one before
 * it), limited against the last point, with a Newton of lo->max_iter
 * iterations (warmOptions' itl2); a lane that fails it, or has no converged
 * last point, solves the point cold by the whole ladder (the plain rung at
 * opts->max_iter, then the stepping rungs). x_out[n_points][n][n_lanes]
 * (NaN where the point did not converge), converged[n_points][n_lanes].
 * The source keeps the last value. */
espice_status espice_lanes_dc_sweep(espice_lanes *l, uint32_t kind, uint32_t instance, const double *values,
    uint32_t n_points, const espice_newton_opts *opts, const espice_ladder_opts *lo, double *x_out,
    uint32_t *converged, double *kernel_ms);

/* ---- AC lanes (src/analysis/ac/ac.zig:40-86, freq.Stream) -------------
 * The reference's AC excitation (frontend/builder.zig:802-817, 1146, 1158,
 * 3124-3151): source (kind, instance)'s AC spec added into drive[2n] (real
 * parts, then imaginary): a voltage source's mag*e^(j phase) on its branch
 * row, a current source's on its two node rows (minus on p, plus on n);
 * phase in degrees. */
espice_status espice_lanes_template_ac_drive(const espice_lanes_template *t, uint32_t kind, uint32_t instance,
    double mag, double phase_deg, double *drive);
/* Every lane's small-signal response at its x (its operating point, left
 * by a Newton or uploaded): the lane's G and C planes from one evaluation
 * there, then for each omega[k] (rad/s) the complex system
 * (G + j omega C) X = drive solved in FP64 (dense LU with partial pivoting
 * on the lane's wavefront). out[n_freq][2][n][n_lanes]: for each
 * frequency the real parts of every unknown, then the imaginary parts;
 * NaN where the lane's matrix is singular. AcChargesUnsupported for
 * templates whose mos1 devices carry charge; lanes up to 64 unknowns. */
espice_status espice_lanes_ac(espice_lanes *l, const double *omega, uint32_t n_freq, const double *drive,
    double *out, double *kernel_ms);

/* ---- the lead-tape lane LU (lane_device.zig refactor :101, solveNeg :170)
 * Builds a LaneTab from one factor of (col_ptr, row_idx, vals) on the host
 * with the template solver's threshold pivoting (lanes.zig:590-607 ownTape,
 * kernels/lane.zig:778-802 buildTab). idx receives *len words (capacity
 * idx_capacity; BUFFER_TOO_SMALL sets *len and copies nothing). FAILED when
 * the matrix is singular. */
espice_status espice_lanes_tab_build(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const double *vals,
    espice_lane_tab *tab, uint32_t *idx, size_t idx_capacity, size_t *len);
/* Replays the tape on every lane on the GPU: a[nnz * n_lanes] and
 * rhs[n * n_lanes] lane-major in; dx[n * n_lanes] = -A_l^-1 rhs_l and
 * fail[n_lanes] (1 = the lane must peel: dead or non-finite pivot, growth,
 * non-finite factor entry, multiplier past 1/pivot_tol, void slot) out. */
espice_status espice_lanes_tab_refactor_solve(const espice_lane_tab *tab, const uint32_t *idx, size_t idx_len,
    uint32_t n_lanes, const double *a, const double *rhs, double *dx, uint8_t *fail, double *kernel_ms);

/* The same replay allocating nothing: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. a, rhs,
 * dx and fail are DEVICE pointers with espice_lanes_tab_refactor_solve's
 * layouts; idx (idx_len words) is on the host and is checked there, then
 * copied into the library's fixed scratch array together with every lane's
 * factor planes: ceil(idx_len / 2) + n_lanes * (max(nnz_l, 1) + max(nnz_u, 1)
 * + 3 n) doubles must fit its 64 MiB, else INVALID_ARGUMENT
 * "ScratchTooSmall". The work is enqueued on `stream` (a hipStream_t; NULL
 * the default stream) and the call returns without waiting: the caller
 * synchronizes before reading dx and fail. One scratch array per GPU,
 * shared by every call on it: calls that overlap on one GPU race. */
espice_status espice_lanes_tab_refactor_solve_device(const espice_lane_tab *tab, const uint32_t *idx, size_t idx_len,
    uint32_t n_lanes, const double *a, const double *rhs, double *dx, uint8_t *fail, void *stream);

/* ---- transient lanes (the device step control of tran.runGpuLanes,
 * src/analysis/tran/tran.zig:1736-1912, with src/solver/kernels/lane.zig's
 * step machine :353-700 and the .meas passes of src/output/scan.zig) ----
 * `.tran dt_init t_stop` with the step ceiling dt_max resolved by the
 * caller (a deck's is min(dt_init, t_stop/50), src/core/query.zig:673-680);
 * method 0 BE, 1 trapezoidal (default), 2 gear-2; tolerances as
 * src/core/numerics.zig:378-399 (reltol 1e-3, abstol 1e-12, vntol 1e-6,
 * chgtol 1e-14, trtol 7); max_iter as optionsFromTolerances (100). */
typedef struct {
    double t_stop, dt_init, dt_max, dt_min, xmu;
    double reltol, abstol, vntol, chgtol, trtol;
    uint32_t method, max_iter, max_steps, record_capacity;
    /* 0: the operating point at t = 0 starts cold (MODEINITJCT seeds,
     * INITFIX); 1: it starts warm at each lane's x (espice_lanes_upload_x
     * or a warm start), as the reference's warm operating point
     * (src/analysis/dc/op.zig:588, 671). */
    uint32_t op_warm;
    /* 0: every Newton tick's linear step in FP64; k > 0: in FP32 with FP64
     * iterative refinement, at most k FP32 solves per tick
     * (espice_lanes_newton_mixed's step; templates of the hand-written
     * kinds or with PSP103, MixedUnsupported otherwise). */
    uint32_t mixed_refine;
    /* The points options' measure handling (include/espice.h:286-301,
     * src/analysis/tran/tran.zig:1059-1064), checked at every accepted
     * point after the passes read it: early_stop 1 ends a lane once every
     * pass has stopped (its values are final); abort_after (NaN off) ends a
     * lane with status 4 once t passes it while a pass whose bit is set in
     * abort_mask (pass k: bit k) is still running. */
    uint32_t early_stop, pad2;
    double abort_after;
    uint64_t abort_mask;
    /* The operating point's reltol at t = 0 when positive (else reltol):
     * a tighter operating point for comparisons between start points, the
     * transient's steps keep reltol. */
    double op_reltol;
} espice_tran_opts;

/* One .meas pass over an unknown `row` (a node voltage or a branch
 * current), scan.zig's kinds: 0 Cross (WHEN/TRIG/TARG: val, td, from, to,
 * rise/fall/cross counts, -1 not given, -2 LAST), 1 At (FIND, op 1 DERIV:
 * at), 2 Extremum (op 0 MIN, 1 MAX, 2 AVG: from, to), 3 Integ (op 0 INTEG,
 * 1 RMS: from, to), 4 Find (FIND, op 1 DERIV, ... WHEN; src/output/
 * scan.zig:339-380): the value of its row where the pass just before it, a
 * Cross without td, from or a LAST count (its WHEN), finds its event;
 * INVALID_ARGUMENT "InvalidMeasure" otherwise. */
typedef struct {
    uint32_t kind, row;
    int32_t rise, fall, cross;
    uint32_t op;
    double val, td, from, to, at;
} espice_meas_pass;

/* Cross: value = the event time; At, Find: value; Extremum: value, aux0 = where
 * (AVG: the last time read); Integ: value, aux0 = from, aux1 = to. NaN when
 * the pass found nothing. */
typedef struct {
    double value, aux0, aux1;
} espice_meas_result;

/* status 0 reached t_stop, 1 dt underflow, 2 no operating point, 3
 * singular operating point, 4 aborted by abort_after; points = recorded
 * rows including t = 0; with mixed_refine, the FP32 solves taken and the
 * Newton ticks handed back to the FP64 factor (0 otherwise). */
typedef struct {
    uint32_t status, points, steps, attempts, rej_newton, rej_lte, bp_landings, nr_iters;
    uint32_t fp32_solves, fp64_fallbacks;
} espice_tran_result;

/* Every lane's transient in one launch, its operating point solved first
 * (cold, or from download/upload_x's state with opts->op_warm):
 * out[n_lanes]; meas[n_lanes * n_passes] lane-major by pass; when
 * record_capacity > 0, times[n_lanes * record_capacity] and
 * rows[n_lanes * record_capacity * n] receive each lane's first
 * record_capacity accepted rows (t = 0 included). download_x then returns
 * each lane's state at its last accepted point. FAILED "LaneTooLarge" for a
 * template whose transient lane does not fit one workgroup's LDS or has
 * more than 64 unknowns; FAILED "TranChargesUnsupported" for a template
 * whose mos1 devices carry charge (TOX, overlap or junction capacitance):
 * the mos1 charge model is not staged. */
espice_status espice_lanes_tran(espice_lanes *l, const espice_tran_opts *opts, const espice_meas_pass *passes,
    uint32_t n_passes, espice_tran_result *out, espice_meas_result *meas, double *times, double *rows,
    double *kernel_ms);

/* ---- Monte Carlo draws (src/frontend/variants.zig:1716-1720, 1862-1895;
 * src/frontend/netlist.zig:683-688; src/core/numerics.zig:416-432) ----
 * Fills out[l] with trial first_trial + l's draw of distribution `fn`
 * (ESPICE_MC_*) at `site`: args = (nom, width[, sigmas][, mult]) as the
 * reference's agauss/gauss/aunif/unif/limit read them. Draws depend only
 * on (seed, trial, site), so lanes may be drawn in any order and count. */
enum { ESPICE_MC_AGAUSS = 0, ESPICE_MC_GAUSS = 1, ESPICE_MC_AUNIF = 2, ESPICE_MC_UNIF = 3, ESPICE_MC_LIMIT = 4 };
espice_status espice_lanes_mc_draw(uint64_t seed, uint64_t site, uint32_t first_trial, uint32_t count, uint32_t fn,
    const double *args, uint32_t n_args, double *out);
/* netlist.siteKey: owner mixed with salt (splitmix64's finalizer). */
uint64_t espice_lanes_site_key(uint64_t owner, uint64_t salt);

#ifdef __cplusplus
}
#endif
#endif
