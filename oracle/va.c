/* Compact models compiled from Verilog-A (egspice_amd.vagen output in
 * egspice_amd/csrc/gen/), compiled here for the CPU with the oracle's own
 * elementary functions. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The generated C is the same text the lane engine compiles; what the
 * oracle adds around it is its own: parameter binding, the collapse of
 * internal unknowns, pattern and stamping (circuit.c), Newton, LU and
 * the transient, and the limiter's passes. Pinned by (tests/test_va.py):
 * central finite differences of every G and C entry, KCL and charge
 * conservation over the terminals, an independent Python evaluation of the
 * compiler's IR (values to a few ulps), and for bsim4va's limiter the
 * hand-written mos1 limiter of the same group semantics. Value parity with the reference's VerA-generated code: parity
 * unpinned (VerA is not in the reference tree and no reference run is
 * possible here). */
#include <math.h>
#include <string.h>

#include "oracle.h"
#include "oracle_internal.h"

#define VA_FN static
#define VA_TABLE static const
#define VA_HOST_TABLE static const
#define VA_RESTRICT restrict
#define VA_EXP om_exp
#define VA_LOG om_log
#define VA_LOG10(x) (om_log(x) * 0.43429448190325182765)
#define VA_POW om_pow
#define VA_SQRT sqrt
#define VA_FABS fabs
#define VA_FLOOR floor
#define VA_CEIL ceil
#define VA_FMA fma
#define VA_INF INFINITY
#define VA_IDIV(a, b) ((b) ? (a) / (b) : 0)
#define VA_IMOD(a, b) ((b) ? (a) % (b) : 0)
#define VA_RTOI(x) ((int)round(x))

#define VA_ATAN atan
#define VA_FETLIM or_fetlim
#define VA_LIMVDS or_limvds
#define VA_PNJLIM or_pnjlim
#define VA_SNAP or_snap_limited

#include "../egspice_amd/csrc/gen/psp103.h"
#include "../egspice_amd/csrc/gen/bsim4va.h"

/* $simparam("gmin"): the deck's gmin, 1e-12 S by default
 * (src/core/numerics.zig:386). */

/* One compiled kind's generated tables and functions. */
typedef struct {
    int nu, nport, nparam, nprep, nsite, ncoll, coll_base, nalias;
    unsigned limw;
    const unsigned *gmask, *cmask;
    const int *collapse, *alias_target;
    const char *const *names, *const *alias_names;
    int (*setup)(const double *, double, double, double *);
    void (*eval)(const double *, const double *, double *, double *);
    void (*eval_tran)(const double *, const double *, double *, double *, double *, double *, double *);
    void (*eval_charge)(const double *, const double *, double *, double *);
    int (*limit)(const double *, const double *, const double *, double *);
    void (*seed)(const double *, const double *, double *);
    void (*limit_args)(const double *, const double *, double *);
} va_kind;

static const va_kind *kind_of(int kind)
{
    static const va_kind psp = {PSP103_NU, PSP103_NPORT, PSP103_NPARAM, PSP103_NPREP, PSP103_NSITE, PSP103_NCOLL,
                                PSP103_COLL_BASE, PSP103_NALIAS, PSP103_LIMW, psp103_gmask, psp103_cmask,
                                psp103_collapse, psp103_alias_target, psp103_param_names, psp103_alias_names,
                                psp103_setup, psp103_eval, psp103_eval_tran, psp103_eval_charge, 0, 0, 0};
    static const va_kind b4 = {BSIM4VA_NU, BSIM4VA_NPORT, BSIM4VA_NPARAM, BSIM4VA_NPREP, BSIM4VA_NSITE,
                               BSIM4VA_NCOLL, BSIM4VA_COLL_BASE, BSIM4VA_NALIAS, BSIM4VA_LIMW, bsim4va_gmask,
                               bsim4va_cmask, bsim4va_collapse, bsim4va_alias_target, bsim4va_param_names,
                               bsim4va_alias_names, bsim4va_setup, bsim4va_eval, bsim4va_eval_tran,
                               bsim4va_eval_charge, bsim4va_limit, bsim4va_seed, bsim4va_limit_args};
    return kind == OR_PSP103 ? &psp : kind == OR_BSIM4 ? &b4 : 0;
}

int or_va_nu(int kind) { return kind_of(kind) ? kind_of(kind)->nu : 0; }
int or_va_nport(int kind) { return kind_of(kind) ? kind_of(kind)->nport : 0; }
int or_va_nparam(int kind) { return kind_of(kind) ? kind_of(kind)->nparam : 0; }
int or_va_nprep(int kind) { return kind_of(kind) ? kind_of(kind)->nprep : 0; }
int or_va_nsite(int kind) { return kind_of(kind) ? kind_of(kind)->nsite : 0; }
unsigned or_va_limw(int kind) { return kind_of(kind) ? kind_of(kind)->limw : 0u; }
unsigned or_va_gmask(int kind, int ru) { return kind_of(kind) ? kind_of(kind)->gmask[ru] : 0u; }
unsigned or_va_cmask(int kind, int ru) { return kind_of(kind) ? kind_of(kind)->cmask[ru] : 0u; }

int or_va_param_index(int kind, const char *name)
{
    const va_kind *K = kind_of(kind);
    if (!K) return -1;
    for (int k = 0; k < K->nparam; k++)
        if (strcasecmp(K->names[k], name) == 0) return k;
    for (int k = 0; k < K->nalias; k++)
        if (strcasecmp(K->alias_names[k], name) == 0) return K->alias_target[k];
    return -1;
}

int or_va_setup(int kind, const double *u, double *p)
{
    const va_kind *K = kind_of(kind);
    return K ? K->setup(u, OR_TEMPERATURE_K, OR_SIMPARAM_GMIN, p) : -1;
}

/* The setup with $simparam("gmin") at `gmin`: the reference reruns it when
 * the operating point's stepping rungs publish theirs
 * (src/device/eval.zig:2596-2612 setHomotopy). */
int or_va_setup_g(int kind, const double *u, double gmin, double *p)
{
    const va_kind *K = kind_of(kind);
    return K ? K->setup(u, OR_TEMPERATURE_K, gmin, p) : -1;
}

/* Rows of one instance's unknowns: ports as given; an internal unknown in a
 * collapsed set takes the set's smallest unknown (a port if it holds one),
 * else a fresh row (rows[u] = -1); -2 - r: unknown r's row. */
int or_va_collapse(int kind, const double *u, const uint32_t *ports, int64_t *rows)
{
    const va_kind *K = kind_of(kind);
    if (!K) return -1;
    const int nu = K->nu, np = K->nport;
    double p[4096];
    int rc = or_va_setup(kind, u, p);
    if (rc != 0) return rc;
    int parent[32];
    for (int k = 0; k < nu; k++) parent[k] = k;
    for (int k = 0; k < K->ncoll; k++) {
        if (p[K->coll_base + k] == 0.0) continue;
        int a = K->collapse[2 * k], b = K->collapse[2 * k + 1];
        while (parent[a] != a) a = parent[a];
        while (parent[b] != b) b = parent[b];
        if (a == b) continue;
        if (a < np && b < np) return -2;
        if (a < b) parent[b] = a;
        else parent[a] = b;
    }
    for (int k = 0; k < nu; k++) {
        int r = k;
        while (parent[r] != r) r = parent[r];
        rows[k] = r < np ? (int64_t)ports[r] : (r == k ? -1 : -2 - r);
    }
    return 0;
}

void or_va_eval(int kind, const double *p, const double *x, double *res, double *jac)
{
    if (kind_of(kind)) kind_of(kind)->eval(p, x, res, jac);
}

void or_va_eval_tran(int kind, const double *p, const double *x, double *res, double *jac, double *q, double *cjac,
                     double *site)
{
    if (kind_of(kind)) kind_of(kind)->eval_tran(p, x, res, jac, q, cjac, site);
}

/* The transient's charge rows and sites alone, values only (what the
 * lane engine evaluates when it re-reads the charges at a solution); the
 * same value operations as or_va_eval_tran, so the same bits. */
void or_va_eval_charge(int kind, const double *p, const double *x, double *q, double *site)
{
    if (kind_of(kind)) kind_of(kind)->eval_charge(p, x, q, site);
}

/* The compiled limiter (the $limit group of the model): the limited image
 * of the unknowns from cur and the previous image or iterate old; returns
 * pnjlim's check. A kind without $limit writes nothing. */
int or_va_limit(int kind, const double *p, const double *cur, const double *old, double *lim)
{
    const va_kind *K = kind_of(kind);
    return K && K->limit ? K->limit(p, cur, old, lim) : 0;
}

/* The MODEINITJCT image of a cold operating point. */
void or_va_seed(int kind, const double *p, const double *cur, double *lim)
{
    const va_kind *K = kind_of(kind);
    if (K && K->seed) K->seed(p, cur, lim);
}

/* The limiter's arguments at x, in call order (test access). */
void or_va_limit_args(int kind, const double *p, const double *x, double *la)
{
    const va_kind *K = kind_of(kind);
    if (K && K->limit_args) K->limit_args(p, x, la);
}
