/* Compact models compiled from Verilog-A (egspice_amd.vagen output in
 * egspice_amd/csrc/gen/), compiled here for the CPU with the oracle's own
 * elementary functions. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The generated C is the same text the lane engine compiles; what the
 * oracle adds around it is its own: parameter binding, the collapse of
 * internal unknowns, pattern and stamping (circuit.c), Newton, LU and
 * the transient. Pinned by (tests/test_va.py): central finite differences
 * of every G and C entry, KCL and charge conservation over the terminals,
 * and an independent Python evaluation of the compiler's IR (values to a
 * few ulps). Value parity with the reference's VerA-generated code: parity
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

#include "../egspice_amd/csrc/gen/psp103.h"

/* $simparam("gmin"): the deck's gmin, 1e-12 S by default
 * (src/core/numerics.zig:386). */
#define OR_SIMPARAM_GMIN 1e-12

int or_va_nu(int kind) { return kind == OR_PSP103 ? PSP103_NU : 0; }
int or_va_nport(int kind) { return kind == OR_PSP103 ? PSP103_NPORT : 0; }
int or_va_nparam(int kind) { return kind == OR_PSP103 ? PSP103_NPARAM : 0; }
int or_va_nprep(int kind) { return kind == OR_PSP103 ? PSP103_NPREP : 0; }
int or_va_nsite(int kind) { return kind == OR_PSP103 ? PSP103_NSITE : 0; }
unsigned or_va_gmask(int kind, int ru) { return kind == OR_PSP103 ? psp103_gmask[ru] : 0u; }
unsigned or_va_cmask(int kind, int ru) { return kind == OR_PSP103 ? psp103_cmask[ru] : 0u; }

int or_va_param_index(int kind, const char *name)
{
    if (kind != OR_PSP103) return -1;
    for (int k = 0; k < PSP103_NPARAM; k++)
        if (strcasecmp(psp103_param_names[k], name) == 0) return k;
    for (int k = 0; k < PSP103_NALIAS; k++)
        if (strcasecmp(psp103_alias_names[k], name) == 0) return psp103_alias_target[k];
    return -1;
}

int or_va_setup(int kind, const double *u, double *p)
{
    if (kind == OR_PSP103) return psp103_setup(u, OR_TEMPERATURE_K, OR_SIMPARAM_GMIN, p);
    return -1;
}

/* Rows of one instance's unknowns: ports as given; an internal unknown in a
 * collapsed set takes the set's smallest unknown (a port if it holds one),
 * else a fresh row (rows[u] = -1); -2 - r: unknown r's row. */
int or_va_collapse(int kind, const double *u, const uint32_t *ports, int64_t *rows)
{
    const int nu = or_va_nu(kind), np = or_va_nport(kind);
    double p[PSP103_NPREP];
    int rc = or_va_setup(kind, u, p);
    if (rc != 0) return rc;
    int parent[32];
    for (int k = 0; k < nu; k++) parent[k] = k;
    for (int k = 0; k < PSP103_NCOLL; k++) {
        if (p[PSP103_COLL_BASE + k] == 0.0) continue;
        int a = psp103_collapse[2 * k], b = psp103_collapse[2 * k + 1];
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
    if (kind == OR_PSP103) psp103_eval(p, x, res, jac);
}

void or_va_eval_tran(int kind, const double *p, const double *x, double *res, double *jac, double *q, double *cjac,
                     double *site)
{
    if (kind == OR_PSP103) psp103_eval_tran(p, x, res, jac, q, cjac, site);
}

/* The transient's charge rows and sites alone, values only (what the
 * lane engine evaluates when it re-reads the charges at a solution); the
 * same value operations as or_va_eval_tran, so the same bits. */
void or_va_eval_charge(int kind, const double *p, const double *x, double *q, double *site)
{
    if (kind == OR_PSP103) psp103_eval_charge(p, x, q, site);
}
