/* CPU oracle (test infrastructure only: the lane engine never links it) of
 * the operating point's ladder after the plain rung, one lane at a time:
 * src/analysis/dc/op.zig:140-330 (solveLadder, afterPlain), its gmin
 * placement :711-790 (gminStamps, ngspice's spMNA_Preorder and LoadGmin),
 * gminFactor :383-387 and gminStalled :352-366, sourcesOffGmin :393-415.
 * Each solve is the lane's own ordered direct Newton (or_newton_hom); the
 * compiled kinds rerun their setup at each solve's $simparam("gmin"), as
 * the reference's setHomotopy does. The JFNK and OPtran rungs are not
 * restated: a lane the final Newton leaves unconverged reports so. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_circuit.h"
#include "oracle_internal.h"

#define GMIN_DECK 1e-12 /* numerics.zig:386 */
#define MAX_LADDER_SOLVES 1000u
#define STALL_SOLVES 20u

/* 10^(1/20), the stall factor (op.zig:346), computed as the reference
 * does: pow of the f64 quotient */
static double stall_factor(void) { return pow(10.0, 1.0 / (double)STALL_SOLVES); }

double or_gmin_factor(double factor, uint32_t iterations, uint32_t budget)
{
    if (iterations <= budget / 4) return fmin(factor * sqrt(factor), 10.0);
    if (iterations > 3 * (budget / 4)) return fmax(sqrt(factor), 1.00005);
    return factor;
}

int or_gmin_stalled(const double *gmin, const int *conv, uint32_t count, double start, double next)
{
    if (count < STALL_SOLVES) return 0;
    const double sf = stall_factor();
    double good = start;
    for (uint32_t i = 0; i < count; i++) {
        if (conv[i]) good = gmin[i];
        if (i < count - STALL_SOLVES) continue;
        const double following = i + 1 < count ? gmin[i + 1] : next;
        if (good / following >= sf) return 0;
    }
    return 1;
}

/* The gmin descent's state between solves (op.zig:185-252). */
typedef struct {
    double factor, good_gmin, gmin_val, start;
    int have_good;
    uint32_t solves, count;
    double steps_g[MAX_LADDER_SOLVES];
    int steps_c[MAX_LADDER_SOLVES];
} descent;

static void descent_init(descent *d, double gmin_start)
{
    d->factor = 10.0;
    d->good_gmin = gmin_start;
    d->start = gmin_start;
    d->gmin_val = gmin_start / 10.0;
    d->have_good = 0;
    d->solves = 0;
    d->count = 0;
}

enum { D_NEXT = 0, D_CLEAN = 1, D_LEAVE = 2 };

/* After the solve at d->gmin_val (tried): D_NEXT with the next gmin in
 * d->gmin_val (*restore: 1 back to the last good x, 2 a cold start, 0 as
 * it is; *save: the solution is the new last good x), D_CLEAN when the
 * target is reached (the clean solve follows), D_LEAVE for source
 * stepping. */
static int descent_after(descent *d, int conv, int refused, uint32_t it, uint32_t budget, int *save, int *restore)
{
    const double tried = d->gmin_val;
    *save = 0;
    *restore = 0;
    if (conv) {
        if (d->gmin_val <= GMIN_DECK) return D_CLEAN;
        *save = 1;
        d->have_good = 1;
        d->good_gmin = d->gmin_val;
        d->factor = or_gmin_factor(d->factor, it, budget);
        if (d->gmin_val < d->factor * GMIN_DECK) {
            d->factor = d->gmin_val / GMIN_DECK;
            d->gmin_val = GMIN_DECK;
        } else {
            d->gmin_val /= d->factor;
        }
    } else {
        if (d->factor < 1.00005) return D_LEAVE;
        d->factor = refused ? sqrt(d->factor) : sqrt(sqrt(d->factor));
        d->gmin_val = d->good_gmin / d->factor;
        *restore = d->have_good ? 1 : 2;
    }
    d->steps_g[d->count] = tried;
    d->steps_c[d->count] = conv;
    d->count++;
    if (or_gmin_stalled(d->steps_g, d->steps_c, d->count, d->start, d->gmin_val)) return D_LEAVE;
    if (++d->solves >= MAX_LADDER_SOLVES) return D_LEAVE;
    return D_NEXT;
}

/* Replays a recorded descent: the solve outcomes conv[k] and it[k]
 * (it = 0 with conv = 0 is a refused factor), from gmin_start; next[k]
 * receives the gmin the state machine tries after solve k and action[k]
 * what it does then. Returns the solves replayed (it stops at the first
 * action that is not D_NEXT). */
uint32_t or_gmin_replay(const int *conv, const uint32_t *it, uint32_t count, double gmin_start, uint32_t budget,
                        double *next, int *action)
{
    descent *d = malloc(sizeof *d);
    descent_init(d, gmin_start);
    uint32_t k = 0;
    for (; k < count; k++) {
        int save, restore;
        action[k] = descent_after(d, conv[k], !conv[k] && it[k] == 0, it[k], budget, &save, &restore);
        next[k] = d->gmin_val;
        if (action[k] != D_NEXT) {
            k++;
            break;
        }
    }
    free(d);
    return k;
}

/* ---------------------------------------------------------- gmin stamps */

/* The first +-1 entry of position j's column whose transpose is +-1 too,
 * and how many such (counting stops at 2): CountTwins of sputils.c. */
static uint32_t twins(const or_circuit *c, const double *g, const uint32_t *perm, uint32_t j, uint32_t *row)
{
    uint32_t cnt = 0;
    const uint32_t col = perm[j];
    for (uint32_t p = or_ckt_col_ptr(c)[col]; p < or_ckt_col_ptr(c)[col + 1]; p++) {
        const uint32_t r = or_ckt_row_idx(c)[p];
        if (r == 0 || fabs(g[p]) != 1.0) continue;
        const uint32_t qs = or_find_slot(c, j, perm[r]);
        if (qs == OR_NONE || fabs(g[qs]) != 1.0) continue;
        if (++cnt == 1) *row = r;
        else break;
    }
    return cnt;
}

static void swap_cols(uint32_t *perm, uint32_t *dcol, uint32_t j, uint32_t r)
{
    dcol[j] = perm[r];
    dcol[r] = perm[j];
    const uint32_t t = perm[j];
    perm[j] = perm[r];
    perm[r] = t;
}

/* gminStamps on lane l's planes at x = 0: |value| == 1 read off an
 * assembly, as the reference reads the plain rung's last one. stamp: 2n
 * words (row i's column or OR_NONE, then its slot). */
void or_gmin_stamps(or_lanes *L, uint32_t l, uint32_t *stamp)
{
    const or_circuit *c = L->c;
    const uint32_t n = or_ckt_n(c), nnz = or_ckt_nnz(c);
    double *g = calloc(nnz + 1, 8), *rhs = calloc(n + 1, 8), *x = calloc(n ? n : 1, 8);
    or_eval_planes(L, l, x, 0, g, rhs, NULL);
    uint32_t *perm = malloc(4 * (n ? n : 1)), *dcol = malloc(4 * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) {
        perm[i] = i;
        dcol[i] = (or_ckt_cur_row(c)[i] && g[or_ckt_diag(c)[i]] == 0.0) ? OR_NONE : i;
    }
    /* spMNA_Preorder: lone twins first, then the first multi-twin column */
    uint32_t start = 1;
    for (;;) {
        int swapped = 0, again = 0;
        for (uint32_t j = start; j < n; j++) {
            if (dcol[j] != OR_NONE) continue;
            uint32_t row = 0;
            const uint32_t cnt = twins(c, g, perm, j, &row);
            if (cnt == 1) {
                swap_cols(perm, dcol, j, row);
                swapped = 1;
            } else if (cnt > 1 && !again) {
                again = 1;
                start = j;
            }
        }
        if (!again) break;
        for (uint32_t j = start; !swapped && j < n; j++) {
            if (dcol[j] != OR_NONE) continue;
            uint32_t row = 0;
            if (twins(c, g, perm, j, &row) > 0) {
                swap_cols(perm, dcol, j, row);
                swapped = 1;
            }
        }
        if (!swapped) break;
    }
    for (uint32_t i = 0; i < 2 * n; i++) stamp[i] = OR_NONE;
    for (uint32_t i = 1; i < n; i++) {
        if (dcol[i] == OR_NONE) continue;
        stamp[i] = dcol[i];
        stamp[n + i] = or_find_slot(c, i, dcol[i]);
    }
    free(g); free(rhs); free(x); free(perm); free(dcol);
}

/* ---------------------------------------------------------------- ladder */

typedef struct {
    or_lanes *L;
    uint32_t l, n, budget;
    const or_newton_opts *plain;
    const uint32_t *stamp;
    double model_gmin;
    uint32_t total;
} lane_ctx;

enum { S_FAIL = 0, S_CONV = 1, S_SINGULAR = 2 };

/* One solve: the shunt at the stamps, the models' gmin, the sources'
 * factor, cold or warm; the stepping budget, or the plain rung's
 * (itl1) when plain_budget. */
static int solve(lane_ctx *k, double *x, double gshunt, double gmodel, double srcfact, int cold, int plain_budget,
                 uint32_t *it)
{
    if (gmodel != k->model_gmin) {
        or_lanes_set_model_gmin(k->L, k->l, gmodel);
        k->model_gmin = gmodel;
    }
    or_newton_opts o = *k->plain;
    o.gmin = gshunt;
    o.init_fix = cold ? 1 : 0;
    o.max_iter = plain_budget ? k->plain->max_iter : k->budget;
    or_outcome r;
    or_newton_hom(k->L, k->l, &o, x, &r, k->stamp, srcfact);
    const int singular = (r.status & 2u) != 0;
    *it = singular ? 0 : r.iterations;
    k->total = k->total + *it > 65535u ? 65535u : k->total + *it;
    return singular ? S_SINGULAR : (r.status & 1u) ? S_CONV : S_FAIL;
}

/* op.zig:393-415: the sources-off point under a diagonal gmin from 1e10
 * times the deck's down to the deck's, a decade per solve, each from the
 * last; true when all eleven converge. */
static int sources_off_gmin(lane_ctx *k, double *x, double lambda)
{
    uint32_t it;
    memset(x, 0, 8 * k->n);
    double g = GMIN_DECK * 1e10;
    for (int i = 0; i < 11; i++) {
        if (solve(k, x, g, GMIN_DECK, lambda, i == 0, 0, &it) != S_CONV) return 0;
        g /= 10;
    }
    return 1;
}

/* Rungs 2-3 and the final Newton for one lane whose plain rung failed;
 * returns the rung that converged (2 gmin stepping, 3 source stepping) or
 * 0. */
static int after_plain(lane_ctx *k, double *x, double gmin_start)
{
    const uint32_t n = k->n;
    double *x_good = calloc(n ? n : 1, 8);
    descent *d = malloc(sizeof *d);
    uint32_t it;
    int rung = 0;
    /* rung 2: dynamic gmin stepping (op.zig:178-255) */
    memset(x, 0, 8 * n);
    descent_init(d, gmin_start);
    for (;;) {
        const int st = solve(k, x, d->gmin_val, d->gmin_val, 1.0, !d->have_good, 0, &it);
        int save, restore;
        const int act = descent_after(d, st == S_CONV, st == S_SINGULAR, it, k->budget, &save, &restore);
        if (act == D_CLEAN) {
            /* ngspice's dynamic_gmin drops the shunt for the last solve */
            if (solve(k, x, 0.0, GMIN_DECK, 1.0, 0, 1, &it) == S_CONV) rung = 2;
            break;
        }
        if (act == D_LEAVE) break;
        if (save) memcpy(x_good, x, 8 * n);
        if (restore == 1) memcpy(x, x_good, 8 * n);
        else if (restore == 2) memset(x, 0, 8 * n);
    }
    free(d);
    if (rung) {
        free(x_good);
        return rung;
    }
    /* rung 3: source stepping (op.zig:257-311) */
    memset(x, 0, 8 * n);
    double lambda = 0.0, lambda_good = -1.0, delta = 0.25;
    for (uint32_t solves = 0; solves < MAX_LADDER_SOLVES; solves++) {
        int conv = solve(k, x, 0.0, GMIN_DECK, lambda, lambda_good < 0.0, 0, &it) == S_CONV;
        if (!conv && lambda_good < 0.0) conv = sources_off_gmin(k, x, lambda);
        if (conv) {
            if (lambda >= 1.0) break;
            lambda_good = lambda;
            memcpy(x_good, x, 8 * n);
            delta *= 1.5;
            lambda = fmin(lambda + delta, 1.0);
        } else {
            delta *= 0.5;
            if (delta < 1e-4) break;
            if (lambda_good < 0.0) break;
            memcpy(x, x_good, 8 * n);
            lambda = fmin(lambda_good + delta, 1.0);
        }
    }
    /* the final Newton at full sources (op.zig:313-320) */
    if (solve(k, x, 0.0, GMIN_DECK, 1.0, 0, 1, &it) == S_CONV) rung = 3;
    free(x_good);
    return rung;
}

/* Every lane's operating point: the plain rung as `plain` says (from x,
 * n * N node-major), then for each lane it leaves unconverged or
 * singular the stepping rungs. out[l]: status 1 converged / 0 not,
 * iterations summed (saturating at 65535); rung[l] 1 plain, 2 gmin
 * stepping, 3 source stepping, 0 none. */
int or_op_ladder(or_lanes *L, const or_newton_opts *plain, double gmin_start, uint32_t budget, double *x,
                 or_outcome *out, uint32_t *rung)
{
    const or_circuit *c = L->c;
    const uint32_t n = or_ckt_n(c), N = L->N;
    double *xl = malloc(8 * (n ? n : 1));
    uint32_t *stamp = malloc(8 * (n ? n : 1));
    for (uint32_t l = 0; l < N; l++) {
        for (uint32_t i = 0; i < n; i++) xl[i] = x[(size_t)i * N + l];
        lane_ctx k = {L, l, n, budget, plain, NULL, OR_SIMPARAM_GMIN, 0};
        /* rung 1, gmin on every diagonal as the caller's options say */
        or_outcome r;
        if (or_newton_hom(L, l, plain, xl, &r, NULL, 1.0) != 0) {
            free(xl);
            free(stamp);
            return -1;
        }
        const int singular = (r.status & 2u) != 0;
        k.total = singular ? 0 : (r.iterations > 65535u ? 65535u : r.iterations);
        int rg = 1;
        if (!((r.status & 1u) && !singular)) {
            or_gmin_stamps(L, l, stamp);
            k.stamp = stamp;
            rg = after_plain(&k, xl, gmin_start);
            if (k.model_gmin != OR_SIMPARAM_GMIN) or_lanes_set_model_gmin(L, l, OR_SIMPARAM_GMIN);
        }
        out[l].status = rg ? 1u : 0u;
        out[l].iterations = k.total;
        if (rung) rung[l] = (uint32_t)rg;
        for (uint32_t i = 0; i < n; i++) x[(size_t)i * N + l] = xl[i];
    }
    free(xl);
    free(stamp);
    return 0;
}

/* ------------------------------------------------------------- dc sweep */

int or_dc_sweep_lanes(or_lanes *L, int kind, uint32_t inst, const double *values, uint32_t npts,
                      const or_newton_opts *plain, double gmin_start, uint32_t budget, double *x_out,
                      uint32_t *converged)
{
    const or_circuit *c = L->c;
    const uint32_t n = or_ckt_n(c), N = L->N;
    const int pk = or_prepared_count(kind);
    double *sol = malloc(8 * (n ? n : 1)), *prev = malloc(8 * (n ? n : 1));
    uint32_t *stamp = malloc(8 * (n ? n : 1));
    for (uint32_t l = 0; l < N; l++) {
        memset(sol, 0, 8 * n);
        memset(prev, 0, 8 * n);
        or_gmin_stamps(L, l, stamp);
        lane_ctx k = {L, l, n, budget, plain, stamp, OR_SIMPARAM_GMIN, 0};
        int cold = 1;
        for (uint32_t pt = 0; pt < npts; pt++) {
            /* the swept value, user and prepared (the dc value once given) */
            L->uprm[kind][((size_t)inst * (kind == OR_VSRC ? OR_VSRC_K : OR_ISRC_K) + OR_VSRC_DC) * N + l] = values[pt];
            L->pprm[kind][((size_t)inst * pk + SRCP_DC) * N + l] = values[pt];
            /* MODEINITPRED: the last two points extrapolated */
            for (uint32_t i = 0; i < n; i++) {
                const double last = sol[i];
                sol[i] = 2.0 * last - prev[i];
                prev[i] = last;
            }
            int done = 0;
            or_outcome r;
            if (!cold) {
                /* prepareWarm, then the warm Newton (warmOptions: itl2) */
                or_limit_pass(L, l, sol, prev, 0);
                or_newton_opts w = *plain;
                w.max_iter = budget;
                w.gmin = 0.0;
                w.init_fix = 0;
                or_newton_hom_lim(L, l, &w, sol, &r, NULL, 1.0, 1);
                done = (r.status & 1u) && !(r.status & 2u);
            }
            if (!done) {
                /* coldStart + the whole ladder */
                memset(sol, 0, 8 * n);
                or_newton_opts p0 = *plain;
                p0.gmin = 0.0;
                p0.init_fix = 1;
                or_newton_hom(L, l, &p0, sol, &r, NULL, 1.0);
                done = (r.status & 1u) && !(r.status & 2u);
                if (!done) {
                    done = after_plain(&k, sol, gmin_start) != 0;
                    if (k.model_gmin != OR_SIMPARAM_GMIN) {
                        or_lanes_set_model_gmin(L, l, OR_SIMPARAM_GMIN);
                        k.model_gmin = OR_SIMPARAM_GMIN;
                    }
                }
            }
            converged[(size_t)pt * N + l] = (uint32_t)done;
            for (uint32_t i = 0; i < n; i++) x_out[((size_t)pt * n + i) * N + l] = done ? sol[i] : NAN;
            cold = !done;
            /* the first point's solution stands for the one before it too */
            if (done && pt == 0) memcpy(prev, sol, 8 * n);
        }
    }
    free(sol);
    free(prev);
    free(stamp);
    return 0;
}
