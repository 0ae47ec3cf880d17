/* Transient analysis of one lane: the oracle of the lane step machine.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates src/analysis/tran/tran.zig's Sim for the shapes the lane
 * engine takes (order <= 2: BE, trapezoidal, gear-2; no uic, VPI, delay
 * lines, state flips or tangents):
 *   tr_init      Sim.init tran.zig:430-705 (charge rings seeded with q(op),
 *                ceiling, delmin/min_break, first dt, dt_old = dt_max)
 *   tr_begin     Sim.begin :749-804 + HostOps.attempt :1167-1190 (predictor)
 *   tr_end       Sim.end :806-1087 (Newton verdict, LTE bound and retry,
 *                order promotion, advanceCurrent and ring rotation,
 *                breakpoint landing, next dt) with floorRetry :1089-1105
 *   tr_newton    converger.newton under integrator.Hook: companion residual
 *                rhs += ag0 (q - q1) [- ag1 i_prev | - ag2 (q1 - q2)],
 *                matrix G + ag0 C (integrator.zig companionAt, combineGC)
 *   step_bound   integrator.zig stepBound/minDel :454-509 over the per-site
 *                charge tape (LTE per ddt site, tran.zig:468-476)
 *   next_bp      tran.zig:1273-1284 nextBp over the sources' breakpoint
 *                caches (device/eval.zig:1592-1598 srcBreak and the batch
 *                cache of nextBreakpointFn :2709-2730)
 * Newton options: converger.optionsFromTolerances(tol, itl4) (max_iter
 * max(itl4, 100) = 100, gmin 0) with residual_tol = inf (tran.zig:313-317).
 * Parity pinned by: tests/test_oracle_tran.py (the reference's printed
 * .meas values and point counts of docs/examples/measure.sp and
 * rc_lowpass.sp, and the RC closed forms).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_internal.h"

#define TR_NH 4 /* charge ring slots: max(4, top order + 2) */

typedef struct {
    int method; /* 0 BE, 1 trap, 2 gear-2 */
    double ag0, ag1, ag2;
} tr_coeffs;

/* integrator.zig coeffs :19-31. */
static tr_coeffs coeffs(int method, double dt, double dt_prev, double xmu)
{
    tr_coeffs c = {method, 0, 1, 0};
    if (method == 0) {
        c.ag0 = 1.0 / dt;
    } else if (method == 1) {
        c.ag0 = 1.0 / dt / (1.0 - xmu);
        c.ag1 = xmu / (1.0 - xmu);
    } else {
        const double r = dt / dt_prev;
        c.ag0 = (1.0 + 2.0 * r) / ((1.0 + r) * dt);
        c.ag2 = r * r / ((1.0 + r) * dt);
    }
    return c;
}

/* The coefficients, exported for the known-value pins of
 * src/analysis/tests/transient.zig:605-636 (ag0, ag1, ag2 into out). */
void or_tran_coeffs(int method, double dt, double dt_prev, double xmu, double *out)
{
    const tr_coeffs c = coeffs(method, dt, dt_prev, xmu);
    out[0] = c.ag0;
    out[1] = c.ag1;
    out[2] = c.ag2;
}

/* tran.zig methodAt :1298: order 1 is BE, else the top method. */
static int method_at(int top, int k) { return k == 1 ? 0 : top; }

/* tran.zig almostEqualUlps :2315 and atStop :1307. */
static int almost_equal_ulps(double a, double b, int64_t max_ulps)
{
    if (a == b) return 1;
    int64_t ia, ib;
    memcpy(&ia, &a, 8);
    memcpy(&ib, &b, 8);
    __int128 la = ia < 0 ? (__int128)INT64_MIN - ia : ia;
    __int128 lb = ib < 0 ? (__int128)INT64_MIN - ib : ib;
    __int128 d = la - lb;
    if (d < 0) d = -d;
    return d <= max_ulps;
}
static int at_stop(double t, double t_stop) { return almost_equal_ulps(t, t_stop, 100); }

static double sel_max(double a, double b) { return a > b ? a : b; } /* maxNoNan's select */

/* integrator.zig lteCoeff :161-171 for order <= 2. */
static double lte_coeff(int method) { return method == 0 ? 0.5 : method == 1 ? 0.08333333333 : 0.2222222222; }

typedef struct {
    double dt, dt1, dt2, reltol, abstol, chgtol, trtol;
} lte_in;

/* stepBound / minDel (integrator.zig:454-509, kernels/lane.zig:257-297):
 * the smallest del over the sites, rooted for order 2; NaN when any read
 * entry is non-finite or a del is 0/0. */
static double step_bound(int method, int cur_method, const double *q0, const double *q1, const double *q2, const double *q3,
                         const double *ip, uint32_t len, tr_coeffs c, lte_in e)
{
    const double coeff = lte_coeff(method);
    const double sum01 = e.dt + e.dt1, sum12 = e.dt1 + e.dt2, sum012 = e.dt + (e.dt1 + e.dt2);
    double vmin = INFINITY, poison = 0;
    for (uint32_t j = 0; j < len; j++) {
        const double qc = q0[j], qp = q1[j], qp2 = q2[j], i_p = ip[j];
        double i_new;
        if (cur_method == 1) i_new = c.ag0 * (qc - qp) - c.ag1 * i_p;
        else if (cur_method == 2) i_new = c.ag0 * (qc - qp) - c.ag2 * (qp - qp2);
        else i_new = c.ag0 * (qc - qp);
        const double volttol = e.abstol + e.reltol * sel_max(fabs(i_new), fabs(i_p));
        const double chargetol = e.reltol * sel_max(sel_max(fabs(qc), fabs(qp)), e.chgtol) / e.dt;
        const double tol = sel_max(volttol, chargetol);
        poison += (qc - qc) + (qp - qp) + (qp2 - qp2) + (i_p - i_p);
        const double f12 = (qp - qp2) / e.dt1;
        double dd = ((qc - qp) / e.dt - f12) / sum01;
        if (method != 0) {
            const double qp3 = q3[j];
            poison += qp3 - qp3;
            const double f123 = (f12 - (qp2 - qp3) / e.dt2) / sum12;
            dd = (dd - f123) / sum012;
        }
        const double del = e.trtol * tol / sel_max(e.abstol, coeff * fabs(dd));
        poison += del == del ? 0 : del;
        vmin = del < vmin ? del : vmin;
    }
    if (poison != poison) return NAN;
    return method == 0 ? vmin : sqrt(vmin);
}

/* integrator.zig companionAt: out = (acc ? out : 0) + ag0 (q0 - q1), then
 * the method's history term. */
static void companion(int method, int accumulate, double *out, const double *q0, const double *q1, const double *q2,
                      const double *i_prev, tr_coeffs c, uint32_t len)
{
    for (uint32_t j = 0; j < len; j++) {
        const double d = c.ag0 * (q0[j] - q1[j]);
        const double base = accumulate ? out[j] + d : d;
        if (method == 1) out[j] = base - c.ag1 * i_prev[j];
        else if (method == 2) out[j] = base - c.ag2 * (q1[j] - q2[j]);
        else out[j] = base;
    }
}

/* ------------------------------------------------------------ breakpoints */

typedef struct {
    double req, at; /* srcBreak's per-model cache */
} src_cache;

typedef struct {
    const or_lanes *L;
    uint32_t l;
    uint32_t n_src;
    double (*prm)[SRCP_K]; /* resolved prepared constants per V source */
    src_cache *cache;
    double lo, hi;         /* the batch cache of nextBreakpointFn */
} tr_bps;

/* The V-source batch's next breakpoint strictly after t, the accepted time
 * being t_acc (device/eval.zig:2709-2730 with srcBreak :1592). */
static double batch_next(tr_bps *b, double t_acc, double t)
{
    if (b->lo <= t && t < b->hi) return b->hi;
    double best = INFINITY;
    for (uint32_t i = 0; i < b->n_src; i++) {
        src_cache *c = &b->cache[i];
        if (t < c->req || t >= c->at) {
            c->req = t;
            c->at = or_src_break(b->prm[i], t_acc, t);
        }
        best = c->at < best ? c->at : best;
    }
    b->lo = t;
    b->hi = best;
    return best;
}

/* ------------------------------------------------------------------ Newton */

typedef struct {
    int on;
    tr_coeffs c;
    const double *q1, *q2, *i_prev;
} tr_hook;

typedef struct {
    or_lanes *L;
    uint32_t l, n, nnz;
    const uint32_t *diag;
    const uint8_t *cur_row;
    or_solver *solver;
    double *g, *a, *rhs, *q, *cc, *dx, *x_old;
} tr_ws;

/* converger.newton (converger.zig:155-319) for one lane at time t, under
 * the hook when `h->on`; x holds the start and returns x_k of the last
 * accepted iterate. Returns 1 converged, 0 not, -1 singular. */
static int tr_newton(tr_ws *w, double *x, double t, const tr_hook *h, const or_newton_opts *o, int init_fix,
                     int lim_active, uint32_t *iterations)
{
    const uint32_t n = w->n, nnz = w->nnz;
    or_evctx ec = {1, t, w->q, h->on ? w->cc : NULL, NULL, NULL};
    uint32_t bits = init_fix ? 2u : 0u;
    double limit = 1e9;
    memcpy(w->x_old, x, 8 * n);
    *iterations = o->max_iter;
    for (uint32_t tick = 0; tick < o->max_iter; tick++) {
        or_eval_planes(w->L, w->l, x, lim_active, w->g, w->rhs, &ec);
        const double *vals = w->g;
        if (h->on) {
            companion(h->c.method, 1, w->rhs, w->q, h->q1, h->q2, h->i_prev, h->c, n);
            for (uint32_t p = 0; p <= nnz; p++) w->a[p] = w->g[p] + h->c.ag0 * w->cc[p];
            vals = w->a;
        }
        if (tick == 0) limit = or_lane_plausible_limit(n, x, w->rhs, vals, w->diag, 1, 0);
        if (or_solver_factor(w->solver, vals) != 0) {
            *iterations = 0;
            return -1;
        }
        or_solver_solve_neg(w->solver, w->rhs, w->dx);
        const double scaled = or_lane_update_norm(n, x, w->dx, w->x_old, w->cur_row, o->reltol, o->abstol, o->vntol, 1, 0);
        const int limited = or_limit_pass(w->L, w->l, x, w->x_old, lim_active);
        lim_active = 1;
        int ok = !limited && tick != 0 && !(scaled >= 1.0);
        if (ok) ok = or_lane_plausible(n, x, limit, 1, 0);
        if (ok) ok = or_lane_residual(n, o->reltol, o->vntol, o->residual_tol, x, w->rhs, vals, w->diag, 1, 0);
        if (ok && (bits & 2u)) {
            bits &= ~2u;
        } else if (ok) {
            memcpy(x, w->x_old, 8 * n);
            *iterations = tick + 1;
            return 1;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------- Sim */

typedef struct {
    double *t;
    double *x;
    uint32_t len, cap, n;
} tr_rec;

static int rec_push(tr_rec *r, double t, const double *x)
{
    if (r->len == r->cap) {
        uint32_t cap = r->cap ? 2 * r->cap : 1024;
        double *nt = realloc(r->t, 8 * (size_t)cap), *nx = realloc(r->x, 8 * (size_t)cap * r->n);
        if (!nt || !nx) {
            if (nt) r->t = nt;
            if (nx) r->x = nx;
            return -1;
        }
        r->t = nt;
        r->x = nx;
        r->cap = cap;
    }
    r->t[r->len] = t;
    memcpy(r->x + (size_t)r->len * r->n, x, 8 * r->n);
    r->len++;
    return 0;
}

int or_tran_run(or_lanes *L, uint32_t l, const or_tran_opts *opt, double **times, double **rows, or_tran_stats *st)
{
    return or_tran_run_from(L, l, opt, NULL, times, rows, st);
}

/* x0 == NULL: the operating point starts cold (seeds, INITFIX); else warm
 * at x0 (n values), without seeds or INITFIX, as op.zig's warm start
 * (src/analysis/dc/op.zig:588, 671). */
int or_tran_run_from(or_lanes *L, uint32_t l, const or_tran_opts *opt, const double *x0, double **times, double **rows,
                     or_tran_stats *st)
{
    const or_circuit *c = L->c;
    const uint32_t n = or_ckt_n(c), nnz = or_ckt_nnz(c), nq = n + 1, n_qt = or_tape_len(c);
    /* the independent sources, voltage then current: both carry the
     * waveform block and its breakpoints (vsource.va, isource.va) */
    const uint32_t n_vsrc = or_ckt_count(c, OR_VSRC), n_src = n_vsrc + or_ckt_count(c, OR_ISRC);
    const int has_charge = n_qt > 0;
    memset(st, 0, sizeof *st);
    *times = NULL;
    *rows = NULL;

    /* per-run copy of the sources with the PULSE defaults of the .tran
     * card filled in (frontend/builder.zig:845-877 resolvePulseDefaults) */
    tr_bps bps = {L, l, n_src, calloc(n_src + 1, sizeof(double[SRCP_K])), calloc(n_src + 1, sizeof(src_cache)), INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n_src; i++) {
        double *p = bps.prm[i];
        or_load_p(L, i < n_vsrc ? OR_VSRC : OR_ISRC, i < n_vsrc ? i : i - n_vsrc, l, p);
        if (p[SRCP_WAVE] == 5.0) { /* builder.zig:854-857 */
            if (p[SRCP_FFM] == -1.0) p[SRCP_FFM] = 5.0 / opt->t_stop;
            if (p[SRCP_FFC] == 0.0) p[SRCP_FFC] = 500.0 / opt->t_stop;
        }
        if (p[SRCP_WAVE] == 2.0 && p[SRCP_SFREQ] == 0) p[SRCP_SFREQ] = 1.0 / opt->t_stop;
        if (p[SRCP_WAVE] == 3.0) { /* builder.zig:862-867 */
            if (p[SRCP_ETD1] == 0) p[SRCP_ETD1] = opt->dt_init;
            if (p[SRCP_ETAU1] == 0) p[SRCP_ETAU1] = opt->dt_init;
            if (p[SRCP_ETD2] == 0) p[SRCP_ETD2] = p[SRCP_ETD1] + opt->dt_init;
            if (p[SRCP_ETAU2] == 0) p[SRCP_ETAU2] = opt->dt_init;
        }
        if (p[SRCP_WAVE] != 1.0) {
            p[SRCP_TD] = 1e30;
        } else {
            if (p[SRCP_TR] <= 0) p[SRCP_TR] = opt->dt_init;
            if (p[SRCP_TF] <= 0) p[SRCP_TF] = opt->dt_init;
            if (p[SRCP_PW] <= 0) p[SRCP_PW] = opt->t_stop;
            if (p[SRCP_PER] <= 0) p[SRCP_PER] = opt->t_stop;
        }
        bps.cache[i].req = INFINITY;
        bps.cache[i].at = INFINITY;
    }
    /* the lane's sources evaluate from the resolved copy */
    double *saved = malloc(8 * ((size_t)n_src * SRCP_K + 1));
    for (uint32_t i = 0; i < n_src; i++)
        for (int k = 0; k < SRCP_K; k++) {
            const uint32_t id = i < n_vsrc ? i : i - n_vsrc;
            double *slot = &L->pprm[i < n_vsrc ? OR_VSRC : OR_ISRC][((size_t)id * SRCP_K + k) * L->N + l];
            saved[i * SRCP_K + k] = *slot;
            *slot = bps.prm[i][k];
        }

    tr_ws w = {L, l, n, nnz, or_ckt_diag(c), or_ckt_cur_row(c), or_solver_new(n, or_ckt_col_ptr(c), or_ckt_row_idx(c)),
               calloc(nnz + 1, 8), calloc(nnz + 1, 8), calloc(nq, 8), calloc(nq, 8), calloc(nnz + 1, 8), calloc(n, 8),
               calloc(n, 8)};
    double *cur = calloc(n, 8), *prev = calloc(n, 8), *trial = calloc(n, 8);
    double *ring = calloc((size_t)TR_NH * nq + 1, 8), *i_prev = calloc(nq, 8);
    double *tring = calloc((size_t)TR_NH * n_qt + 1, 8), *ti_prev = calloc(n_qt + 1, 8);
    double *qh[TR_NH], *th[TR_NH];
    for (int k = 0; k < TR_NH; k++) {
        qh[k] = ring + (size_t)k * nq;
        th[k] = tring + (size_t)k * n_qt;
    }
    tr_rec rec = {NULL, NULL, 0, 0, n};
    int rc = 0;

    /* the transient's operating point: sources at t = 0, a cold start, or
     * warm at x0 */
    or_newton_opts nopt = {opt->reltol, opt->abstol, opt->vntol, opt->residual_tol_op, 0.0, opt->max_iter, x0 ? 0u : 1u};
    {
        tr_hook off = {0};
        uint32_t it;
        for (uint32_t i = 0; i < n; i++) cur[i] = x0 ? x0[i] : 0.0;
        if (!x0) or_seed_pass(L, l, cur);
        if (tr_newton(&w, cur, 0.0, &off, &nopt, x0 ? 0 : 1, x0 ? 0 : 1, &it) != 1) {
            st->status = 2;
            goto done;
        }
    }
    memcpy(prev, cur, 8 * n);
    if (has_charge) {
        /* q(op) seeds a flat history; i_prev starts at 0 */
        or_evctx ec = {1, 0.0, qh[1], NULL, th[1], NULL};
        or_eval_planes(L, l, cur, 0, w.g, w.rhs, &ec);
        for (int k = 2; k < TR_NH; k++) {
            memcpy(qh[k], qh[1], 8 * nq);
            memcpy(th[k], th[1], 8 * n_qt);
        }
    }
    const double dt_max = opt->dt_max;
    const double delmin = 1e-11 * dt_max, min_break = 10.0 * delmin;
    if (rec_push(&rec, 0.0, cur)) {
        st->status = 3;
        goto done;
    }
    double dt = fmin(fmin(opt->dt_init, opt->t_stop / 100.0) / 10.0, dt_max);
    {
        const double bp0 = batch_next(&bps, 0.0, min_break);
        if (bp0 != INFINITY) dt = fmin(dt, 0.1 * bp0);
    }
    dt /= 10.0;
    double t = 0.0, dt_old[6];
    for (int k = 0; k < 6; k++) dt_old[k] = dt_max;
    int order = 1, run_order = 1;
    uint32_t order_run = 0, steps = 0;
    double bp_next = INFINITY, bp_save_dt = opt->t_stop / 50.0;
    int bp_clamped = 0;
    nopt.residual_tol = INFINITY;
    nopt.init_fix = 0;

    for (;;) {
        /* begin */
        if (!(t < opt->t_stop && !at_stop(t, opt->t_stop) && steps < opt->max_steps)) break;
        const double dt_att = dt, t_dev = t + dt;
        const int method = method_at(opt->method, order);
        const tr_coeffs cf = coeffs(method, dt, dt_old[0], opt->xmu);
        const double xfact = dt / dt_old[0];
        for (uint32_t i = 0; i < n; i++) trial[i] = cur[i] + xfact * (cur[i] - prev[i]);
        /* MODEINITPRED: the predicted iterate limited against the last
         * accepted point, with the limiters cleared by the previous solve
         * (HostOps.attempt tran.zig:1185-1189; applyLimits eval.zig:2430) */
        (void)or_limit_pass(L, l, trial, cur, 0);
        tr_hook hk = {1, cf, qh[1], qh[2], i_prev};
        uint32_t iters = 0;
        const int nr = tr_newton(&w, trial, t_dev, &hk, &nopt, 0, 1, &iters);
        /* end */
        st->attempts++;
        st->nr_iters += nr == -1 ? 0 : iters;
        if (nr != 1) {
            st->rej_newton++;
            order = 1;
            dt = dt / 8.0;
            goto floor_retry;
        }
        double dt_next = steps == 0 ? dt : fmin(dt * 2.0, dt_max);
        const lte_in e = {dt, dt_old[0], dt_old[1], opt->reltol, opt->abstol, opt->chgtol, opt->trtol};
        if (has_charge) {
            /* the charge re-read at the solution, unlimited (HostOps.charges
             * tran.zig:1207-1213, evalQRange eval.zig:1086-1108) */
            or_evctx ec = {1, t_dev, qh[0], NULL, th[0], NULL};
            or_eval_planes(L, l, trial, 0, w.g, w.rhs, &ec);
        }
        if (has_charge && steps > 0) {
            const double del = step_bound(method, method, th[0], th[1], th[2], th[3], ti_prev, n_qt, cf, e);
            if (!(del >= 0.9 * dt)) {
                st->rej_lte++;
                if (del != del) {
                    order = 1;
                    dt = dt / 8.0;
                } else {
                    dt = del;
                }
                goto floor_retry;
            }
            dt_next = fmin(fmin(sel_max(del, opt->dt_min), 2.0 * dt), dt_max);
        }
        if (has_charge) {
            order_run = order == run_order ? order_run + 1 : 1;
            run_order = order;
            if (steps > 0 && order == 1) {
                const int promote = method_at(opt->method, 2);
                const double trial_del = step_bound(promote, method, th[0], th[1], th[2], th[3], ti_prev, n_qt, cf, e);
                if (trial_del == trial_del) {
                    const double nd2 = fmin(2.0 * dt, trial_del);
                    if (nd2 > 1.05 * dt) order = 2;
                    dt_next = fmin(sel_max(nd2, opt->dt_min), dt_max);
                }
            }
            /* advanceCurrent on rows and tape, then [cur, prev, ...] ->
             * [stale, cur, prev, ...] */
            companion(method, 0, i_prev, qh[0], qh[1], qh[2], i_prev, cf, nq);
            companion(method, 0, ti_prev, th[0], th[1], th[2], ti_prev, cf, n_qt);
            double *lq = qh[TR_NH - 1], *lt = th[TR_NH - 1];
            for (int k = TR_NH - 1; k > 0; k--) {
                qh[k] = qh[k - 1];
                th[k] = th[k - 1];
            }
            qh[0] = lq;
            th[0] = lt;
        }
        for (int k = 5; k > 0; k--) dt_old[k] = dt_old[k - 1];
        dt_old[0] = dt;
        {
            double *stale = prev;
            prev = cur;
            cur = trial;
            trial = stale;
        }
        t = t + dt;
        steps++;
        const double t_new = t;
        if (bp_next != INFINITY) {
            const double bp = bp_next;
            const int landed = bp_clamped ? fabs(t_new - bp) <= min_break
                                          : (bp - t_new <= delmin || almost_equal_ulps(t_new, bp, 100));
            if (landed) {
                st->bp_landings++;
                order = 1;
                double shrink = bp_save_dt;
                const double nb = batch_next(&bps, t_new, t_new + min_break);
                if (nb != INFINITY) shrink = fmin(shrink, nb - t_new);
                dt_next = fmin(dt_next, 0.1 * shrink);
            }
            bp_next = INFINITY;
        }
        if (rec_push(&rec, t_new, cur)) {
            st->status = 3;
            goto done;
        }
        bp_next = batch_next(&bps, t_new, t_new + min_break);
        bp_clamped = 0;
        if (bp_next != INFINITY) {
            const double dt_to_bp = bp_next - t_new;
            if (dt_to_bp < dt_next) {
                bp_save_dt = dt_next;
                dt_next = dt_to_bp;
                bp_clamped = 1;
            }
        }
        dt = dt_next;
        if (t_new + dt > opt->t_stop) dt = opt->t_stop - t_new;
        continue;
    floor_retry:
        if (dt < opt->dt_min && dt_att > opt->dt_min && t + opt->dt_min > t) dt = opt->dt_min;
        {
        if (dt < opt->dt_min) {
            st->status = 1;
            break;
        }
    }
    }
    st->steps = steps;
done:
    st->points = rec.len;
    *times = rec.t;
    *rows = rec.x;
    rc = (int)rec.len;
    for (uint32_t i = 0; i < n_src; i++)
        for (int k = 0; k < SRCP_K; k++) {
            const uint32_t id = i < n_vsrc ? i : i - n_vsrc;
            L->pprm[i < n_vsrc ? OR_VSRC : OR_ISRC][((size_t)id * SRCP_K + k) * L->N + l] = saved[i * SRCP_K + k];
        }
    free(saved);
    or_solver_free(w.solver);
    free(w.g); free(w.a); free(w.rhs); free(w.q); free(w.cc); free(w.dx); free(w.x_old);
    free(cur); free(prev); free(trial); free(ring); free(i_prev); free(tring); free(ti_prev);
    free(bps.prm); free(bps.cache);
    return rc;
}
