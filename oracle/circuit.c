/* Circuit template, planes and Newton drivers of the oracle. TEST
 * INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   or_ckt_node/or_ckt_add  frontend/builder.zig:559-565 rowOf (first touch)
 *                           and :238-266 addDevice (internal rows at add)
 *   or_ckt_freeze           device/Circuit.zig:112-201 (diagonals, then each
 *                           batch's structural entries; PatternBuilder.toCsc
 *                           abi.zig:599-643 sorts and dedups per column;
 *                           diag slots; slot tapes eval.zig:514-532, ground
 *                           entries to the trash slot nnz)
 *   eval_planes             analysis/Circuit.zig evalNewton: clear g and
 *                           rhs, every batch in proto order (evalRange,
 *                           eval.zig:780-925), then groundStamp :861-864
 *   limit_pass              eval.zig:1137-1168 limitRange + applyLimits
 *                           :2431-2436 (lim_active after the first call)
 *   newton_lane (mode 0)    the GPU lane Newton: lanes.zig:1516-1601 ticks,
 *                           lane_device.zig refactor :101, peelDense :120,
 *                           solveNeg :170, assemble :282, update :316, gates
 *                           :327, finalize :351 (lane 0's first-step tape,
 *                           lanes.zig:590-607 ownTape)
 *   newton_scalar (mode 1)  converger.zig:155-319 newton, finalizeStep
 *                           :458-526, with direct.zig's Solver per lane
 *   or_dc_sweep             analysis/dc/dc.zig:23-44 run, Sweep.beginPoint
 *                           :453-490 (predictor), prepareWarm :493-497,
 *                           warmOptions :441-443
 * Parity pinned by: tests/test_oracle_newton.py (the diode_iv golden DC
 * rows; mode 0 == mode 1 bitwise where the lead tape's pivots hold; KCL
 * of every converged lane; FD Jacobian).
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_internal.h"

#define NONE OR_NONE

int or_solver_is_tri(const or_solver *s);

/* A diode's unknowns are (a, c, ai); ai is a's own row when RS is 0. */
/* The compiled kinds' entries are filled from va.c at load. */
static int NU[OR_NKIND] = {3, 2, 2, 2, 3, 4, 0};
static int NPORT[OR_NKIND] = {2, 2, 2, 2, 2, 4, 0};
static int UK[OR_NKIND] = {OR_VSRC_K, OR_ISRC_K, OR_RES_K, OR_CAP_K, OR_DIO_K, OR_M1_K, 0};
static int PK[OR_NKIND] = {SRCP_K, SRCP_K, 2, 1, DIOP_K, M1P_K, 0};
static int NSITE[OR_NKIND] = {0, 0, 0, 1, 1, 0, 0};
static int is_va(int kind) { return kind >= OR_PSP103; }
/* room for one instance's user or prepared parameters */
#define OR_PARAM_MAX 128
_Static_assert(OR_VSRC_K <= OR_PARAM_MAX && SRCP_K <= OR_PARAM_MAX && OR_DIO_K <= OR_PARAM_MAX && DIOP_K <= OR_PARAM_MAX &&
                   OR_M1_K <= OR_PARAM_MAX && M1P_K <= OR_PARAM_MAX,
               "every kind's parameters fit OR_PARAM_MAX");

/* Structural pattern of each kind (row-major bit masks over local
 * unknowns): G-plane entries and C-plane entries. */
static unsigned gmask(int kind, int ru)
{
    switch (kind) {
    case OR_VSRC: return ru == 2 ? 0x3u : 0x4u;
    case OR_RES: return 0x3u;
    case OR_DIODE: return ru == 0 ? 0x5u : ru == 1 ? 0x6u : 0x7u;
    case OR_MOS1: return ru == 1 ? 0x0u : 0xfu;
    case OR_PSP103:
    case OR_BSIM4: return or_va_gmask(kind, ru);
    default: return 0;
    }
}
static unsigned cmask(int kind, int ru)
{
    switch (kind) {
    case OR_CAP: return 0x3u;
    case OR_DIODE: return ru == 0 ? 0x0u : 0x6u;
    case OR_MOS1: return 0xfu;
    case OR_PSP103:
    case OR_BSIM4: return or_va_cmask(kind, ru);
    default: (void)ru; return 0;
    }
}
static unsigned pmask(int kind, int ru) { return gmask(kind, ru) | cmask(kind, ru); }
/* Residual rows a kind writes in a static evaluation. */
static unsigned RMASK[OR_NKIND] = {0x7u, 0x3u, 0x3u, 0x0u, 0x7u, 0xdu, 0u};
/* Unknowns each kind's limiter writes. */
static unsigned LIMW[OR_NKIND] = {0, 0, 0, 0, 0x4u, 0xbu, 0, 0};

__attribute__((constructor)) static void va_tables(void)
{
    for (int k = OR_PSP103; k < OR_NKIND; k++) {
        NU[k] = or_va_nu(k);
        NPORT[k] = or_va_nport(k);
        UK[k] = or_va_nparam(k);
        PK[k] = or_va_nprep(k);
        NSITE[k] = or_va_nsite(k);
        RMASK[k] = (1u << NU[k]) - 1u;
        LIMW[k] = or_va_limw(k);
    }
}
/* scratch big enough for any kind's parameters */
#define OR_VA_PARAM_MAX 4096

typedef struct { uint32_t *p; size_t len, cap; } vu32;
typedef struct { double *p; size_t len, cap; } vf64;
static void vpush(vu32 *v, uint32_t x)
{
    if (v->len == v->cap) { v->cap = v->cap ? 2 * v->cap : 64; v->p = realloc(v->p, 4 * v->cap); }
    v->p[v->len++] = x;
}
static void fpush(vf64 *v, double x)
{
    if (v->len == v->cap) { v->cap = v->cap ? 2 * v->cap : 64; v->p = realloc(v->p, 8 * v->cap); }
    v->p[v->len++] = x;
}

struct or_circuit {
    uint32_t n;
    /* net name table (open addressing) */
    char **names;
    uint32_t *name_row;
    size_t name_cap, name_len;
    uint8_t *cur_flag;
    size_t cur_cap;
    vu32 nodes[OR_NKIND];
    vf64 uprm[OR_NKIND];
    uint32_t count[OR_NKIND];
    int order[OR_NKIND];
    int n_order;
    /* frozen */
    int frozen;
    uint32_t nnz;
    uint32_t *col_ptr, *row_idx, *diag;
    uint8_t *cur_row;
    uint32_t *slots[OR_NKIND];
    uint32_t kslots[OR_NKIND];
};

or_circuit *or_ckt_new(void)
{
    or_circuit *c = calloc(1, sizeof(or_circuit));
    c->n = 1; /* ground */
    c->name_cap = 64;
    c->names = calloc(c->name_cap, sizeof(char *));
    c->name_row = calloc(c->name_cap, sizeof(uint32_t));
    c->cur_cap = 64;
    c->cur_flag = calloc(c->cur_cap, 1);
    return c;
}

void or_ckt_free(or_circuit *c)
{
    if (!c) return;
    for (size_t i = 0; i < c->name_cap; i++) free(c->names[i]);
    free(c->names); free(c->name_row); free(c->cur_flag);
    for (int k = 0; k < OR_NKIND; k++) { free(c->nodes[k].p); free(c->uprm[k].p); free(c->slots[k]); }
    free(c->col_ptr); free(c->row_idx); free(c->diag); free(c->cur_row);
    free(c);
}

static uint64_t hash_str(const char *s)
{
    uint64_t h = 1469598103934665603ull;
    for (; *s; s++) { h ^= (unsigned char)*s; h *= 1099511628211ull; }
    return h;
}

static uint32_t add_row(or_circuit *c, int current)
{
    uint32_t r = c->n++;
    if (r >= c->cur_cap) {
        size_t nc = c->cur_cap * 2;
        c->cur_flag = realloc(c->cur_flag, nc);
        memset(c->cur_flag + c->cur_cap, 0, nc - c->cur_cap);
        c->cur_cap = nc;
    }
    c->cur_flag[r] = (uint8_t)current;
    return r;
}

uint32_t or_ckt_node(or_circuit *c, const char *name)
{
    if (strcmp(name, "0") == 0 || strcmp(name, "gnd") == 0) return 0;
    if (2 * (c->name_len + 1) > c->name_cap) {
        size_t oc = c->name_cap, ncap = oc * 2;
        char **on = c->names;
        uint32_t *orow = c->name_row;
        c->names = calloc(ncap, sizeof(char *));
        c->name_row = calloc(ncap, sizeof(uint32_t));
        c->name_cap = ncap;
        for (size_t i = 0; i < oc; i++)
            if (on[i]) {
                size_t h = hash_str(on[i]) & (ncap - 1);
                while (c->names[h]) h = (h + 1) & (ncap - 1);
                c->names[h] = on[i];
                c->name_row[h] = orow[i];
            }
        free(on);
        free(orow);
    }
    size_t h = hash_str(name) & (c->name_cap - 1);
    while (c->names[h]) {
        if (strcmp(c->names[h], name) == 0) return c->name_row[h];
        h = (h + 1) & (c->name_cap - 1);
    }
    c->names[h] = strdup(name);
    c->name_row[h] = add_row(c, 0);
    c->name_len++;
    return c->name_row[h];
}

int32_t or_ckt_add(or_circuit *c, int kind, const uint32_t *ports, const double *params)
{
    if (kind < 0 || kind >= OR_NKIND || c->frozen) return -1;
    int seen = 0;
    for (int i = 0; i < c->n_order; i++) seen |= c->order[i] == kind;
    if (!seen) c->order[c->n_order++] = kind;
    if (is_va(kind)) {
        /* the compiled kind's collapse from its setup on these parameters */
        int64_t rows[32];
        if (!params || or_va_collapse(kind, params, ports, rows) != 0) return -1;
        const size_t base = c->nodes[kind].len;
        for (int u = 0; u < NU[kind]; u++) {
            if (rows[u] >= 0) vpush(&c->nodes[kind], (uint32_t)rows[u]);
            else if (rows[u] == -1) vpush(&c->nodes[kind], add_row(c, 0));
            else vpush(&c->nodes[kind], c->nodes[kind].p[base + (size_t)(-2 - rows[u])]);
        }
        for (int k = 0; k < UK[kind]; k++) fpush(&c->uprm[kind], params[k]);
        return (int32_t)c->count[kind]++;
    }
    for (int p = 0; p < NPORT[kind]; p++) vpush(&c->nodes[kind], ports[p]);
    /* internal rows fresh, in unknown order, except a diode's ai, which is
     * its anode unless RS > 0 (builder.zig:258-263, diode.va:364-371) */
    for (int u = NPORT[kind]; u < NU[kind]; u++) {
        if (kind == OR_DIODE && !(params && params[OR_DIO_RS] > 0.0)) vpush(&c->nodes[kind], ports[0]);
        else vpush(&c->nodes[kind], add_row(c, kind == OR_VSRC));
    }
    for (int k = 0; k < UK[kind]; k++) fpush(&c->uprm[kind], params ? params[k] : NAN);
    return (int32_t)c->count[kind]++;
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

static uint32_t find_slot(const or_circuit *c, uint32_t row, uint32_t col)
{
    uint32_t lo = c->col_ptr[col], hi = c->col_ptr[col + 1];
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (c->row_idx[mid] < row) lo = mid + 1;
        else hi = mid;
    }
    return (lo < c->col_ptr[col + 1] && c->row_idx[lo] == row) ? lo : NONE;
}

uint32_t or_find_slot(const or_circuit *c, uint32_t row, uint32_t col) { return find_slot(c, row, col); }

int or_ckt_freeze(or_circuit *c)
{
    const uint32_t n = c->n;
    size_t nk = n;
    for (int t = 0; t < c->n_order; t++) {
        int kind = c->order[t];
        nk += (size_t)c->count[kind] * NU[kind] * NU[kind];
    }
    uint64_t *keys = malloc(8 * nk);
    size_t m = 0;
    for (uint32_t i = 0; i < n; i++) keys[m++] = ((uint64_t)i << 32) | i;
    for (int t = 0; t < c->n_order; t++) {
        int kind = c->order[t];
        for (uint32_t id = 0; id < c->count[kind]; id++) {
            const uint32_t *nd = c->nodes[kind].p + (size_t)id * NU[kind];
            for (int ru = 0; ru < NU[kind]; ru++)
                for (int cu = 0; cu < NU[kind]; cu++)
                    if ((pmask(kind, ru) >> cu & 1) && nd[ru] != 0 && nd[cu] != 0)
                        keys[m++] = ((uint64_t)nd[cu] << 32) | nd[ru];
        }
    }
    qsort(keys, m, 8, cmp_u64);
    c->col_ptr = calloc(n + 1, 4);
    c->row_idx = malloc(4 * m);
    uint32_t nnz = 0;
    for (size_t i = 0; i < m; i++) {
        if (i > 0 && keys[i] == keys[i - 1]) continue;
        uint32_t col = (uint32_t)(keys[i] >> 32);
        c->row_idx[nnz++] = (uint32_t)keys[i];
        c->col_ptr[col + 1]++;
    }
    for (uint32_t j = 0; j < n; j++) c->col_ptr[j + 1] += c->col_ptr[j];
    free(keys);
    c->nnz = nnz;
    c->diag = malloc(4 * n);
    for (uint32_t i = 0; i < n; i++) c->diag[i] = find_slot(c, i, i);
    c->cur_row = malloc(n);
    memcpy(c->cur_row, c->cur_flag, n);
    for (int kind = 0; kind < OR_NKIND; kind++) {
        uint32_t ks = 0;
        for (int ru = 0; ru < NU[kind]; ru++) ks += (uint32_t)__builtin_popcount(pmask(kind, ru));
        c->kslots[kind] = ks;
        c->slots[kind] = malloc(4 * ((size_t)c->count[kind] * ks + 1));
        size_t k = 0;
        for (uint32_t id = 0; id < c->count[kind]; id++) {
            const uint32_t *nd = c->nodes[kind].p + (size_t)id * NU[kind];
            for (int ru = 0; ru < NU[kind]; ru++)
                for (int cu = 0; cu < NU[kind]; cu++)
                    if (pmask(kind, ru) >> cu & 1)
                        c->slots[kind][k++] = (nd[ru] == 0 || nd[cu] == 0) ? nnz : find_slot(c, nd[ru], nd[cu]);
        }
    }
    c->frozen = 1;
    return 0;
}

uint32_t or_ckt_n(const or_circuit *c) { return c->n; }
uint32_t or_ckt_nnz(const or_circuit *c) { return c->nnz; }
const uint32_t *or_ckt_col_ptr(const or_circuit *c) { return c->col_ptr; }
const uint32_t *or_ckt_row_idx(const or_circuit *c) { return c->row_idx; }
const uint32_t *or_ckt_diag(const or_circuit *c) { return c->diag; }
const uint8_t *or_ckt_cur_row(const or_circuit *c) { return c->cur_row; }
uint32_t or_ckt_count(const or_circuit *c, int kind) { return c->count[kind]; }
int or_prepared_count(int kind) { return PK[kind]; }
uint32_t or_tape_len(const or_circuit *c)
{
    uint32_t k = 0;
    for (int kind = 0; kind < OR_NKIND; kind++) k += c->count[kind] * (uint32_t)NSITE[kind];
    return k;
}

/* ------------------------------------------------------------------ lanes */

/* struct or_lanes lives in oracle_internal.h (shared with tran.c). */

or_lanes *or_lanes_new(const or_circuit *c, uint32_t N)
{
    or_lanes *L = calloc(1, sizeof(or_lanes));
    L->c = c;
    L->N = N;
    for (int kind = 0; kind < OR_NKIND; kind++) {
        size_t cnt = c->count[kind];
        L->uprm[kind] = malloc(8 * (cnt * UK[kind] * N + 1));
        L->pprm[kind] = calloc(cnt * PK[kind] * N + 1, 8);
        L->lim[kind] = calloc(cnt * NU[kind] * N + 1, 8);
        for (size_t i = 0; i < cnt; i++)
            for (int k = 0; k < UK[kind]; k++)
                for (uint32_t l = 0; l < N; l++)
                    L->uprm[kind][(i * UK[kind] + k) * N + l] = c->uprm[kind].p[i * UK[kind] + k];
    }
    return L;
}

void or_lanes_free(or_lanes *L)
{
    if (!L) return;
    for (int k = 0; k < OR_NKIND; k++) { free(L->uprm[k]); free(L->pprm[k]); free(L->lim[k]); }
    free(L);
}

void or_lanes_set_param(or_lanes *L, int kind, uint32_t inst, uint32_t k, const double *values)
{
    for (uint32_t l = 0; l < L->N; l++) L->uprm[kind][((size_t)inst * UK[kind] + k) * L->N + l] = values[l];
}

void or_lanes_prepare(or_lanes *L)
{
    const uint32_t N = L->N;
    for (int kind = 0; kind < OR_NKIND; kind++) {
        uint32_t cnt = L->c->count[kind];
        static double u[OR_VA_PARAM_MAX], p[OR_VA_PARAM_MAX];
        for (uint32_t i = 0; i < cnt; i++)
            for (uint32_t l = 0; l < N; l++) {
                for (int k = 0; k < UK[kind]; k++) u[k] = L->uprm[kind][((size_t)i * UK[kind] + k) * N + l];
                switch (kind) {
                case OR_PSP103:
                case OR_BSIM4:
                    if (or_va_setup(kind, u, p) != 0)
                        for (int k = 0; k < PK[kind]; k++) p[k] = NAN;
                    break;
                case OR_MOS1: or_prep_mos1(u, p); break;
                case OR_DIODE: or_prep_diode(u, p); break;
                case OR_VSRC:
                case OR_ISRC: {
                    /* vsource.va parameter defaults; NaN = not given */
                    /* vsource.va's defaults: 0, the -1 sentinels of TR, TF, PW,
                     * PER and PWL's r=; the point count an integer in the table */
                    for (int k = 0; k < SRCP_K; k++) {
                        const int sentinel = (k >= SRCP_TR && k <= SRCP_PER) || k == SRCP_PWL_R;
                        const double dflt = (sentinel || k == SRCP_FFM) ? -1.0
                                            : (k == SRCP_ETAU1 || k == SRCP_ETAU2) ? 1.0e-9
                                            : k == SRCP_FMDI ? 90.0 : 0.0;
                        p[k] = u[k] == u[k] ? u[k] : dflt;
                    }
                    {
                        const double len = p[SRCP_PWL_LEN];
                        p[SRCP_PWL_LEN] = len >= 1.0 ? (len < OR_PWL_CAP ? floor(len) : OR_PWL_CAP) : 0.0;
                    }
                    /* dc not given: the waveform's t = 0 value, as the builder
                     * fills it (models/vsource.va:216-221) */
                    if (u[OR_VSRC_DC] != u[OR_VSRC_DC]) {
                        if (p[SRCP_WAVE] == 1.0) p[SRCP_DC] = p[SRCP_V1];
                        else if (p[SRCP_WAVE] == 2.0)
                            p[SRCP_DC] = p[SRCP_SVO] + p[SRCP_SVA] * om_sin(2.0 * VA_PI * p[SRCP_SPHASE] / 360.0);
                        else if (p[SRCP_WAVE] == 3.0)
                            p[SRCP_DC] = p[SRCP_EV1];
                        else if (p[SRCP_WAVE] == 4.0)
                            p[SRCP_DC] = p[SRCP_PWL_V];
                    }
                    break;
                }
                case OR_RES: /* resistor.va: rt = r, I = V / rt, slope 1 / rt */
                    p[0] = u[OR_RES_R];
                    p[1] = 1.0 / u[OR_RES_R];
                    break;
                case OR_CAP: p[0] = u[OR_CAP_C]; break;
                }
                for (int k = 0; k < PK[kind]; k++) L->pprm[kind][((size_t)i * PK[kind] + k) * N + l] = p[k];
            }
    }
}

/* Lane l's compiled instances set up again with $simparam("gmin") at
 * `gmin` (Circuit.zig:1394-1405 setHomotopy: the models that read it rerun
 * their setup). */
void or_lanes_set_model_gmin(or_lanes *L, uint32_t l, double gmin)
{
    const uint32_t N = L->N;
    double u[OR_VA_PARAM_MAX], p[OR_VA_PARAM_MAX];
    for (int kind = 0; kind < OR_NKIND; kind++) {
        if (!is_va(kind)) continue;
        for (uint32_t i = 0; i < L->c->count[kind]; i++) {
            for (int k = 0; k < UK[kind]; k++) u[k] = L->uprm[kind][((size_t)i * UK[kind] + k) * N + l];
            if (or_va_setup_g(kind, u, gmin, p) != 0)
                for (int k = 0; k < PK[kind]; k++) p[k] = NAN;
            for (int k = 0; k < PK[kind]; k++) L->pprm[kind][((size_t)i * PK[kind] + k) * N + l] = p[k];
        }
    }
}

/* ------------------------------------------------------------ evaluation */

void or_load_p(const or_lanes *L, int kind, uint32_t id, uint32_t l, double *p)
{
    for (int k = 0; k < PK[kind]; k++) p[k] = L->pprm[kind][((size_t)id * PK[kind] + k) * L->N + l];
}

static double *lim_at(const or_lanes *L, int kind, uint32_t id, uint32_t l)
{
    return L->lim[kind] + ((size_t)l * L->c->count[kind] + id) * NU[kind];
}

/* One compiled instance into the planes, in the same order as the other
 * kinds: its rows, then its G mask's entries row-major (and in a transient
 * its charges, C entries and one tape entry per ddt() site). With limiting
 * active it is evaluated at the limited image of the unknowns its limiter
 * writes, and x minus the image, dotted with each row's structural
 * gradient, is added to the residual and the charge rows (the charge
 * sites are not corrected: only the unlimited re-read at a solution feeds
 * the step control). */
static void eval_va_inst(const or_lanes *L, int kind, uint32_t id, uint32_t l, const uint32_t *nd, const uint32_t *sl,
                         const double *x, int lim_active, double *g, double *rhs, const or_evctx *ec,
                         uint32_t *tape_at)
{
    const or_circuit *c = L->c;
    const int nu = NU[kind];
    double p[OR_VA_PARAM_MAX], xl[32], res[32], jac[1024], q[32], cjac[1024], site[256], corr[32];
    or_load_p(L, kind, id, l, p);
    int live = 0;
    const unsigned writes = LIMW[kind];
    for (int u = 0; u < nu; u++) {
        xl[u] = x[nd[u]];
        corr[u] = 0.0;
    }
    if (writes && lim_active) {
        const double *lm = lim_at(L, kind, id, l);
        for (int u = 0; u < nu; u++) {
            const double xi = xl[u];
            if (writes >> u & 1) xl[u] = lm[u];
            corr[u] = xi - xl[u];
            live |= corr[u] != 0;
        }
    }
    const int charged = ec && (ec->q || ec->cc || ec->tape);
    if (charged) or_va_eval_tran(kind, p, xl, res, jac, q, cjac, site);
    else or_va_eval(kind, p, xl, res, jac);
    if (live) {
        int kg = 0, kc = 0;
        for (int ru = 0; ru < nu; ru++) {
            double acc = 0;
            for (int cu = 0; cu < nu; cu++)
                if (gmask(kind, ru) >> cu & 1) acc += jac[kg++] * corr[cu];
            res[ru] += acc;
        }
        if (charged)
            for (int ru = 0; ru < nu; ru++) {
                double acc = 0;
                for (int cu = 0; cu < nu; cu++)
                    if (cmask(kind, ru) >> cu & 1) acc += cjac[kc++] * corr[cu];
                q[ru] += acc;
            }
    }
    if (ec && ec->tape)
        for (int k = 0; k < NSITE[kind]; k++) ec->tape[(*tape_at)++] = site[k];
    size_t k = 0, kg = 0, kc = 0;
    for (int ru = 0; ru < nu; ru++) {
        uint32_t row = nd[ru] == 0 ? c->n : nd[ru];
        rhs[row] += res[ru];
        if (ec && ec->q && cmask(kind, ru) != 0) ec->q[row] += q[ru];
        for (int cu = 0; cu < nu; cu++) {
            if (!(pmask(kind, ru) >> cu & 1)) continue;
            if (gmask(kind, ru) >> cu & 1) g[sl[k]] += jac[kg++];
            if (cmask(kind, ru) >> cu & 1) {
                if (ec && ec->cc) ec->cc[sl[k]] += cjac[kc];
                kc++;
            }
            k++;
        }
    }
}

/* g (nnz+1 with trash), rhs (n+1 with trash row n) at x. */
void or_eval_planes(const or_lanes *L, uint32_t l, const double *x, int lim_active, double *g, double *rhs,
                    const or_evctx *ec)
{
    const or_circuit *c = L->c;
    memset(g, 0, 8 * (c->nnz + 1));
    memset(rhs, 0, 8 * (c->n + 1));
    if (ec && ec->q) memset(ec->q, 0, 8 * (c->n + 1));
    if (ec && ec->cc) memset(ec->cc, 0, 8 * (c->nnz + 1));
    const double sf = ec && ec->srcfact ? *ec->srcfact : 1.0;
    uint32_t tape_at = 0;
    for (int t = 0; t < c->n_order; t++) {
        int kind = c->order[t];
        const int nu = NU[kind];
        for (uint32_t id = 0; id < c->count[kind]; id++) {
            const uint32_t *nd = c->nodes[kind].p + (size_t)id * nu;
            const uint32_t *sl = c->slots[kind] + (size_t)id * c->kslots[kind];
            if (is_va(kind)) {
                eval_va_inst(L, kind, id, l, nd, sl, x, lim_active, g, rhs, ec, &tape_at);
                continue;
            }
            double p[OR_PARAM_MAX], lx[4], xl[4];
            or_devout o;
            memset(&o, 0, sizeof o);
            or_load_p(L, kind, id, l, p);
            for (int u = 0; u < nu; u++) xl[u] = lx[u] = x[nd[u]];
            int has_lim = kind == OR_MOS1 || kind == OR_DIODE;
            if (has_lim && lim_active) {
                const double *lm = lim_at(L, kind, id, l);
                unsigned writes = LIMW[kind];
                for (int u = 0; u < nu; u++)
                    if (writes >> u & 1) lx[u] = lm[u];
            }
            switch (kind) {
            case OR_VSRC:
                o.res[0] = lx[2];
                o.res[1] = -lx[2];
                o.res[2] = lx[0] - lx[1] - (ec && ec->tran ? or_source_value(p, ec->t) : p[SRCP_DC]) * sf;
                o.jac[0][2] = 1.0;
                o.jac[1][2] = -1.0;
                o.jac[2][0] = 1.0;
                o.jac[2][1] = -1.0;
                break;
            case OR_ISRC: { /* isource.va:216-330: dc at an OP, the waveform in a transient */
                const double out = (ec && ec->tran ? or_source_value(p, ec->t) : p[SRCP_DC]) * sf;
                o.res[0] = out;
                o.res[1] = -out;
                break;
            }
            case OR_RES: {
                double i = (lx[0] - lx[1]) / p[0];
                o.res[0] = i;
                o.res[1] = -i;
                o.jac[0][0] = p[1];
                o.jac[0][1] = -p[1];
                o.jac[1][0] = -p[1];
                o.jac[1][1] = p[1];
                break;
            }
            case OR_CAP: {
                /* capacitor.va: I(p,n) <+ c*ddt(V(p,n)), charge c*V(p,n) */
                double qv = p[0] * (lx[0] - lx[1]);
                o.q[0] = qv;
                o.q[1] = -qv;
                o.cjac[0][0] = p[0];
                o.cjac[0][1] = -p[0];
                o.cjac[1][0] = -p[0];
                o.cjac[1][1] = p[0];
                if (ec && ec->tape) ec->tape[tape_at++] = qv;
                break;
            }
            case OR_DIODE: or_eval_diode(p, lx, &o, ec && ec->tran); break;
            case OR_MOS1: or_eval_mos1(p, lx, &o); break;
            }
            /* limiting correction: F += J (x - x_lim) on every row */
            if (has_lim && lim_active) {
                double corr[4];
                int live = 0;
                for (int u = 0; u < nu; u++) {
                    corr[u] = xl[u] - lx[u];
                    live |= corr[u] != 0;
                }
                if (live)
                    for (int ru = 0; ru < nu; ru++) {
                        double acc = 0;
                        for (int u = 0; u < nu; u++) acc += o.jac[ru][u] * corr[u];
                        o.res[ru] += acc;
                        /* the charges take the same correction (eval.zig:893-897) */
                        if (cmask(kind, ru) != 0) {
                            double qa = 0;
                            for (int u = 0; u < nu; u++) qa += o.cjac[ru][u] * corr[u];
                            o.q[ru] += qa;
                        }
                    }
            }
            /* a diode's LTE site: its junction charge, row ai */
            if (kind == OR_DIODE && ec && ec->tape) ec->tape[tape_at++] = o.q[2];
            size_t k = 0;
            for (int ru = 0; ru < nu; ru++) {
                uint32_t row = nd[ru] == 0 ? c->n : nd[ru];
                if (RMASK[kind] >> ru & 1) rhs[row] += o.res[ru];
                if (ec && ec->q && (cmask(kind, ru) != 0)) ec->q[row] += o.q[ru];
                for (int cu = 0; cu < nu; cu++) {
                    if (!(pmask(kind, ru) >> cu & 1)) continue;
                    if (gmask(kind, ru) >> cu & 1) g[sl[k]] += o.jac[ru][cu];
                    if (ec && ec->cc && (cmask(kind, ru) >> cu & 1)) ec->cc[sl[k]] += o.cjac[ru][cu];
                    k++;
                }
            }
        }
    }
    /* groundStamp */
    g[c->diag[0]] += 1.0;
    rhs[0] += x[0];
}

/* Limit pass on lane l: cur = x, old = previous image or x_old. Returns 1
 * when any instance clamped (not converged). */
int or_limit_pass(or_lanes *L, uint32_t l, const double *x, const double *x_old, int lim_active)
{
    const or_circuit *c = L->c;
    int flag = 0;
    for (int t = 0; t < c->n_order; t++) {
        int kind = c->order[t];
        const unsigned writes = LIMW[kind];
        if (writes == 0) continue;
        const int nu = NU[kind];
        for (uint32_t id = 0; id < c->count[kind]; id++) {
            const uint32_t *nd = c->nodes[kind].p + (size_t)id * nu;
            double p[OR_VA_PARAM_MAX];  /* per call: lanes run on several threads */
            double cur[32], old[32], lim[32];
            double *lm = lim_at(L, kind, id, l);
            or_load_p(L, kind, id, l, p);
            for (int u = 0; u < nu; u++) {
                cur[u] = x[nd[u]];
                old[u] = ((writes >> u & 1) && lim_active) ? lm[u] : x_old[nd[u]];
            }
            int ch = is_va(kind)         ? or_va_limit(kind, p, cur, old, lim)
                     : kind == OR_MOS1 ? or_limit_mos1(p, cur, old, lim)
                                       : or_limit_diode(p, cur, old, lim);
            if (ch) flag = 1;
            for (int u = 0; u < nu; u++)
                if (writes >> u & 1) lm[u] = lim[u];
        }
    }
    return flag;
}

void or_seed_pass(or_lanes *L, uint32_t l, const double *x)
{
    const or_circuit *c = L->c;
    for (int t = 0; t < c->n_order; t++) {
        int kind = c->order[t];
        const unsigned writes = LIMW[kind];
        if (writes == 0) continue;
        const int nu = NU[kind];
        for (uint32_t id = 0; id < c->count[kind]; id++) {
            const uint32_t *nd = c->nodes[kind].p + (size_t)id * nu;
            double p[OR_VA_PARAM_MAX];  /* per call: lanes run on several threads */
            double cur[32], lim[32];
            double *lm = lim_at(L, kind, id, l);
            or_load_p(L, kind, id, l, p);
            for (int u = 0; u < nu; u++) cur[u] = x[nd[u]];
            if (is_va(kind)) or_va_seed(kind, p, cur, lim);
            else if (kind == OR_MOS1) or_seed_mos1(p, cur, lim);
            else or_seed_diode(p, cur, lim);
            for (int u = 0; u < nu; u++)
                if (writes >> u & 1) lm[u] = lim[u];
        }
    }
}

void or_eval_lane(or_lanes *L, uint32_t l, const double *x, double *rhs, double *vals)
{
    const or_circuit *c = L->c;
    double *g = malloc(8 * (c->nnz + 1)), *r = malloc(8 * (c->n + 1));
    or_eval_planes(L, l, x, 0, g, r, NULL);
    memcpy(rhs, r, 8 * c->n);
    memcpy(vals, g, 8 * c->nnz);
    free(g);
    free(r);
}

/* Lane l's G and C planes at x (nnz values each, the ground row pinned in
 * G), from one transient-mode evaluation without limiting: the small-signal
 * linearization of the AC lanes (src/analysis/ac/ac.zig:48, FreqSolver). */
void or_eval_lane_gc(or_lanes *L, uint32_t l, const double *x, double *g_out, double *c_out)
{
    const or_circuit *c = L->c;
    double *g = malloc(8 * (c->nnz + 1)), *r = malloc(8 * (c->n + 1));
    double *q = malloc(8 * (c->n + 1)), *cc = malloc(8 * (c->nnz + 1));
    uint32_t nt = or_tape_len(c);
    double *tape = malloc(8 * (nt + 1));
    or_evctx ec = {1, 0.0, q, cc, tape, NULL};
    or_eval_planes(L, l, x, 0, g, r, &ec);
    memcpy(g_out, g, 8 * c->nnz);
    memcpy(c_out, cc, 8 * c->nnz);
    free(g); free(r); free(q); free(cc); free(tape);
}

/* ---------------------------------------------------------------- newton */

/* Gate bits, kernels/lane.zig:703-716. */
#define G_INITFIX 2u
#define G_CONVERGED 4u
#define G_SINGULAR 8u
#define G_PASSED 16u

/* Shared per-lane Newton over an abstract linear step. `lin` returns 0 and
 * dx on success, nonzero for a singular lane. */
typedef int (*lin_fn)(void *ctx, uint32_t l, const double *a, const double *rhs, double *dx);

/* `stamp` (2n words, or NULL for every diagonal): where gmin lands, row
 * i at column stamp[i] (OR_NONE: nowhere) and slot stamp[n + i]
 * (converger.zig:700-718 loadGmin). `srcfact`: the sources' factor, NULL
 * for 1. */
static void newton_one(or_lanes *L, uint32_t l, const or_newton_opts *o, double *x, or_outcome *out, lin_fn lin, void *lctx, int lim_active,
                       const uint32_t *stamp, const double *srcfact)
{
    const or_circuit *c = L->c;
    const uint32_t n = c->n, nnz = c->nnz;
    double *g = malloc(8 * (nnz + 1)), *rhs = malloc(8 * (n + 1));
    double *dx = malloc(8 * n), *x_old = malloc(8 * n);
    uint32_t bits = o->init_fix ? G_INITFIX : 0;
    double limit = 1e9;
    or_evctx sec = {0, 0.0, NULL, NULL, NULL, srcfact};
    memcpy(x_old, x, 8 * n);
    out->status = 0;
    out->iterations = o->max_iter;
    for (uint32_t tick = 0; tick < o->max_iter; tick++) {
        or_eval_planes(L, l, x, lim_active, g, rhs, srcfact ? &sec : NULL);
        if (o->gmin > 0 && stamp) {
            for (uint32_t i = 0; i < n; i++)
                if (stamp[i] != OR_NONE) g[stamp[n + i]] += o->gmin;
            for (uint32_t i = 0; i < n; i++)
                if (stamp[i] != OR_NONE) rhs[i] += o->gmin * x[stamp[i]];
        } else if (o->gmin > 0) {
            for (uint32_t i = 0; i < n; i++) {
                g[c->diag[i]] += o->gmin;
            }
            for (uint32_t i = 0; i < n; i++) rhs[i] += o->gmin * x[i];
        }
        if (tick == 0) limit = or_lane_plausible_limit(n, x, rhs, g, c->diag, 1, 0);
        if (lin(lctx, l, g, rhs, dx) != 0) {
            out->status = 2;
            out->iterations = 0;
            break;
        }
        double scaled = or_lane_update_norm(n, x, dx, x_old, c->cur_row, o->reltol, o->abstol, o->vntol, 1, 0);
        int limited = or_limit_pass(L, l, x, x_old, lim_active);
        lim_active = 1;
        int ok = !limited && tick != 0 && !(scaled >= 1.0);
        if (ok) ok = or_lane_plausible(n, x, limit, 1, 0);
        if (ok) ok = or_lane_residual(n, o->reltol, o->vntol, o->residual_tol, x, rhs, g, c->diag, 1, 0);
        if (ok && (bits & G_INITFIX)) {
            bits &= ~G_INITFIX;
        } else if (ok) {
            memcpy(x, x_old, 8 * n);
            out->status = 1;
            out->iterations = tick + 1;
            break;
        }
    }
    free(g); free(rhs); free(dx); free(x_old);
}

/* mode 0: replay of the lead tape; a lane whose replay fails (dead or
 * non-finite pivot, growth, wide multiplier, void slot) factors on its own
 * direct.Solver, the host peel of lanes.zig:1603-1653, which the engine
 * runs on the device for every n. */
typedef struct {
    const or_lane_tab *tab;
    const uint32_t *idx;
    double *lx, *ux, *udiag, *w, *y;
    const or_circuit *c;
    or_solver *own;
    uint32_t peels;
} tape_ctx;

static int lin_tape(void *vctx, uint32_t l, const double *a, const double *rhs, double *dx)
{
    (void)l;
    tape_ctx *t = vctx;
    if (or_lane_refactor(t->tab, t->idx, a, t->lx, t->ux, t->udiag, t->w, 1, 0)) {
        if (!t->own) t->own = or_solver_new(t->c->n, t->c->col_ptr, t->c->row_idx);
        t->peels++;
        if (or_solver_factor(t->own, a) != 0) return 1;
        or_solver_solve_neg(t->own, rhs, dx);
        return 0;
    }
    or_lane_solve_neg(t->tab, t->idx, t->lx, t->ux, t->udiag, rhs, t->y, dx, 1, 0);
    return 0;
}

static int lin_direct(void *vctx, uint32_t l, const double *a, const double *rhs, double *dx)
{
    (void)l;
    or_solver *s = vctx;
    if (or_solver_factor(s, a) != 0) return 1;
    or_solver_solve_neg(s, rhs, dx);
    return 0;
}

int or_lead_tab(or_lanes *L, const double *x, int init_fix, or_lane_tab *tab, uint32_t **idx, size_t *len)
{
    const or_circuit *c = L->c;
    const uint32_t N = L->N, n = c->n;
    double *x0 = malloc(8 * n);
    for (uint32_t i = 0; i < n; i++) x0[i] = x[(size_t)i * N];
    if (init_fix) or_seed_pass(L, 0, x0);
    double *g = malloc(8 * (c->nnz + 1)), *rhs = malloc(8 * (n + 1));
    or_eval_planes(L, 0, x0, init_fix, g, rhs, NULL);
    or_solver *s = or_solver_new(n, c->col_ptr, c->row_idx);
    int rc = -1;
    if (s && or_solver_factor(s, g) == 0 && !or_solver_is_tri(s)) {
        *idx = or_build_tab(or_solver_lu(s), c->col_ptr, or_solver_q(s), OR_GROWTH_LIMIT, OR_PIVOT_TOL, tab, len);
        rc = 0;
    }
    or_solver_free(s);
    free(x0); free(g); free(rhs);
    return rc;
}

/* Shared, read-only state of one or_newton_op call; workers take lanes
 * first + k*stride. The reference's CPU path solves each point on its own
 * (op.zig:588,671) with the template's one column order. */
typedef struct {
    or_lanes *L;
    int mode;
    const or_newton_opts *o;
    double *x;
    or_outcome *out;
    or_lane_tab tab;
    uint32_t *idx;
    const uint32_t *q;
    uint64_t diag_fill;
    uint32_t first, stride;
    int rc;
} op_job;

static void *op_worker(void *arg)
{
    op_job *j = arg;
    or_lanes *L = j->L;
    const or_circuit *c = L->c;
    const uint32_t N = L->N, n = c->n;
    const or_newton_opts *o = j->o;
    double *xl = malloc(8 * (n ? n : 1));
    for (uint32_t l = j->first; l < N; l += j->stride) {
        for (uint32_t i = 0; i < n; i++) xl[i] = j->x[(size_t)i * N + l];
        if (o->init_fix) or_seed_pass(L, l, xl);
        if (j->mode == 0) {
            tape_ctx t = {&j->tab, j->idx, NULL, NULL, NULL, NULL, NULL, c, NULL, 0};
            t.lx = calloc(j->tab.nnz_l + 1, 8);
            t.ux = calloc(j->tab.nnz_u + 1, 8);
            t.udiag = calloc(n, 8);
            t.w = calloc(n, 8);
            t.y = calloc(n, 8);
            newton_one(L, l, o, xl, &j->out[l], lin_tape, &t, o->init_fix, NULL, NULL);
            if (t.peels) j->out[l].status |= 4u;
            or_solver_free(t.own);
            free(t.lx); free(t.ux); free(t.udiag); free(t.w); free(t.y);
        } else {
            or_solver *s = or_solver_new_ordered(n, c->col_ptr, c->row_idx, j->q, j->diag_fill);
            if (!s) {
                j->rc = -1;
                break;
            }
            newton_one(L, l, o, xl, &j->out[l], lin_direct, s, o->init_fix, NULL, NULL);
            or_solver_free(s);
        }
        for (uint32_t i = 0; i < n; i++) j->x[(size_t)i * N + l] = xl[i];
    }
    free(xl);
    return NULL;
}

int or_newton_op_mt(or_lanes *L, int mode, const or_newton_opts *o, double *x, or_outcome *out, uint32_t threads)
{
    const or_circuit *c = L->c;
    const uint32_t n = c->n;
    op_job base;
    memset(&base, 0, sizeof base);
    base.L = L;
    base.mode = mode;
    base.o = o;
    base.x = x;
    base.out = out;
    size_t len = 0;
    if (mode == 0 && or_lead_tab(L, x, o->init_fix, &base.tab, &base.idx, &len) != 0) return -1;
    uint32_t *q = malloc(sizeof(uint32_t) * (n ? n : 1));
    if (or_order(n, c->col_ptr, c->row_idx, q) != 0) {
        free(q);
        free(base.idx);
        return -1;
    }
    base.q = q;
    base.diag_fill = or_diag_fill(n, c->col_ptr, c->row_idx, q);
    if (threads < 1) threads = 1;
    if (threads > L->N) threads = L->N ? L->N : 1;
    op_job *jobs = malloc(sizeof(op_job) * threads);
    pthread_t *tid = malloc(sizeof(pthread_t) * threads);
    for (uint32_t t = 0; t < threads; t++) {
        jobs[t] = base;
        jobs[t].first = t;
        jobs[t].stride = threads;
    }
    for (uint32_t t = 1; t < threads; t++) pthread_create(&tid[t], NULL, op_worker, &jobs[t]);
    op_worker(&jobs[0]);
    int rc = jobs[0].rc;
    for (uint32_t t = 1; t < threads; t++) {
        pthread_join(tid[t], NULL);
        if (jobs[t].rc) rc = jobs[t].rc;
    }
    free(jobs);
    free(tid);
    free(q);
    free(base.idx);
    return rc;
}

int or_newton_op(or_lanes *L, int mode, const or_newton_opts *o, double *x, or_outcome *out)
{
    return or_newton_op_mt(L, mode, o, x, out, 1);
}

/* One lane's static Newton under a homotopy (the ladder's stepping
 * solves): gmin o->gmin at `stamp`, sources times srcfact, cold (the
 * limiters' seeds, INITFIX) when o->init_fix, on the lane's own ordered
 * direct solver, as or_newton_op's mode 1. */
int or_newton_hom(or_lanes *L, uint32_t l, const or_newton_opts *o, double *xl, or_outcome *out, const uint32_t *stamp,
                  double srcfact)
{
    return or_newton_hom_lim(L, l, o, xl, out, stamp, srcfact, o->init_fix != 0);
}

/* The same with the limiters' images live from the start when
 * lim_active (a `.dc` point's predictor limited against the last point). */
int or_newton_hom_lim(or_lanes *L, uint32_t l, const or_newton_opts *o, double *xl, or_outcome *out,
                      const uint32_t *stamp, double srcfact, int lim_active)
{
    const or_circuit *c = L->c;
    const uint32_t n = c->n;
    uint32_t *q = malloc(sizeof(uint32_t) * (n ? n : 1));
    if (or_order(n, c->col_ptr, c->row_idx, q) != 0) {
        free(q);
        return -1;
    }
    or_solver *s = or_solver_new_ordered(n, c->col_ptr, c->row_idx, q, or_diag_fill(n, c->col_ptr, c->row_idx, q));
    free(q);
    if (!s) return -1;
    if (o->init_fix) or_seed_pass(L, l, xl);
    newton_one(L, l, o, xl, out, lin_direct, s, lim_active, stamp, &srcfact);
    or_solver_free(s);
    return 0;
}

/* ------------------------------------------------------------- dc sweep */

int or_dc_sweep(or_lanes *L, uint32_t vsrc, double start, double step, uint32_t npts, const or_newton_opts *o, double *x_out, uint32_t *converged)
{
    const or_circuit *c = L->c;
    const uint32_t n = c->n, N = L->N;
    double *sol = calloc(n, 8), *prev = calloc(n, 8);
    or_solver *s = or_solver_new(n, c->col_ptr, c->row_idx);
    if (!s) return -1;
    double v = start;
    or_newton_opts warm = *o;
    warm.init_fix = 0;
    for (uint32_t pt = 0; pt < npts; pt++) {
        if (pt != 0) v += step;
        for (uint32_t l = 0; l < N; l++) L->pprm[OR_VSRC][(size_t)vsrc * N + l] = v;
        for (uint32_t i = 0; i < n; i++) {
            double last = sol[i];
            sol[i] = 2.0 * last - prev[i];
            prev[i] = last;
        }
        or_outcome oc = {0, 0};
        if (pt != 0) {
            /* prepareWarm: limit the predictor against the last point */
            or_limit_pass(L, 0, sol, prev, 0);
            newton_one(L, 0, &warm, sol, &oc, lin_direct, s, 1, NULL, NULL);
        }
        if (!(oc.status & 1)) {
            /* cold point: the plain operating-point rung from zero */
            memset(sol, 0, 8 * n);
            or_newton_opts cold = *o;
            cold.init_fix = 1;
            or_seed_pass(L, 0, sol);
            newton_one(L, 0, &cold, sol, &oc, lin_direct, s, 1, NULL, NULL);
        }
        converged[pt] = oc.status & 1;
        memcpy(x_out + (size_t)pt * n, sol, 8 * n);
        /* the first point's solution stands for the one before it too
         * (dctrcurv.c:457-462, dc.zig:524-527) */
        if (pt == 0 && (oc.status & 1)) memcpy(prev, sol, 8 * n);
    }
    or_solver_free(s);
    free(sol);
    free(prev);
    return 0;
}
