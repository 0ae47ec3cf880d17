/* Internal declarations shared by the oracle's C files. TEST
 * INFRASTRUCTURE ONLY (see oracle.h). */
#ifndef ESPICE_ORACLE_INTERNAL_H
#define ESPICE_ORACLE_INTERNAL_H

#include <stdint.h>

/* $temperature of every run: 27 degC, the SPICE default. */
#define OR_TEMPERATURE_K 300.15

#define DUAL_W 4
typedef struct { double v; double d[DUAL_W]; } dual;

/* One instance's local residual, dF/du, charge and dQ/du. */
typedef struct {
    double res[4];
    double jac[4][4];
    double q[4];
    double cjac[4][4];
} or_devout;

/* Prepared MOS1 constants per instance and lane. */
enum { M1P_TYPE = 0, M1P_VT, M1P_VBIT, M1P_PHIT, M1P_GAMMA, M1P_BETA, M1P_LAMBDA, M1P_ISATS, M1P_ISATD, M1P_GMIN,
       M1P_VCRITS, M1P_VCRITD, M1P_COX, M1P_CZBD, M1P_CZBS, M1P_CZBDSW, M1P_CZBSSW, M1P_PBT, M1P_MJ, M1P_MJSW,
       M1P_FC, M1P_F2D, M1P_F3D, M1P_F4D, M1P_F2S, M1P_F3S, M1P_F4S, M1P_CGSO, M1P_CGDO, M1P_CGBO, M1P_SEED, M1P_K };
/* Prepared diode constants. */
enum { DIOP_IS = 0, DIOP_VTE, DIOP_GMIN, DIOP_VCRIT, DIOP_CJO, DIOP_VJ, DIOP_MT, DIOP_FC, DIOP_TT, DIOP_QON, DIOP_RSON,
       DIOP_RST, DIOP_K };

void or_prep_mos1(const double *u, double *p);
void or_prep_diode(const double *u, double *p);
void or_eval_mos1(const double *p, const double *lx, or_devout *o);
void or_eval_diode(const double *p, const double *lx, or_devout *o, int tran);
int or_limit_mos1(const double *p, const double *cur, const double *old, double *lim);
int or_limit_diode(const double *p, const double *cur, const double *old, double *lim);
void or_seed_mos1(const double *p, const double *cur, double *lim);
void or_seed_diode(const double *p, const double *cur, double *lim);
/* Prepared V/I source constants: the DC value, then the waveform (0 DC,
 * 1 PULSE, 2 SIN), PULSE(V1 V2 TD TR TF PW PER) with the -1 sentinels of
 * models/vsource.va:58-66 for TR, TF, PW and PER not given, and SIN(VO VA
 * FREQ TD THETA PHASE). */
#define OR_PWL_CAP 16 /* PWL points a table holds here (vsource.va: 64) */
enum { SRCP_DC = 0, SRCP_WAVE, SRCP_V1, SRCP_V2, SRCP_TD, SRCP_TR, SRCP_TF, SRCP_PW, SRCP_PER, SRCP_SVO, SRCP_SVA,
       SRCP_SFREQ, SRCP_STD, SRCP_STHETA, SRCP_SPHASE, SRCP_PWL_LEN, SRCP_PWL_R, SRCP_PWL_TD, SRCP_PWL_T,
       SRCP_PWL_V = SRCP_PWL_T + OR_PWL_CAP, SRCP_EV1 = SRCP_PWL_V + OR_PWL_CAP, SRCP_EV2, SRCP_ETD1, SRCP_ETAU1,
       SRCP_ETD2, SRCP_ETAU2, SRCP_FVO, SRCP_FVA, SRCP_FFM, SRCP_FMDI, SRCP_FFC, SRCP_FTD, SRCP_FPHM, SRCP_FPHC,
       SRCP_AVA, SRCP_AVO, SRCP_AMF, SRCP_AFC, SRCP_ATD, SRCP_APHC, SRCP_APHS, SRCP_K };
/* The source's value at transient time t (models/vsource.va:233-247). */
double or_source_value(const double *p, double t);
#define VA_PI 3.14159265358979323846     /* constants.vams M_PI */
#define VA_TWO_PI 6.28318530717958647652 /* M_TWO_PI */
/* ngspice's PULSE breakpoint after tq, requested at the accepted time
 * t_acc (src/device/eval.zig:1636-1650 spicePulseBreak); inf when none. */
double or_pulse_break(const double *p, double t_acc, double tq);
double or_pulse_timer_break(const double *p, double t);
double or_src_break(const double *p, double t_acc, double tq);

/* What one evaluation adds beyond the static planes: the transient's time
 * (sources read their waveform), and the charge outputs. */
typedef struct {
    int tran;
    double t;
    double *q;    /* n + 1 row charges (ground and trash row included), or NULL */
    double *cc;   /* nnz + 1 C-plane values, or NULL */
    double *tape; /* one charge per LTE site, or NULL */
    /* the independent sources' factor, $simparam("sourceScaleFactor")
     * (vsource.va:333, isource.va:332); NULL for 1 */
    const double *srcfact;
} or_evctx;

struct or_lanes {
    const struct or_circuit *c;
    uint32_t N;
    double *uprm[OR_NKIND]; /* [(inst*UK + k)*N + l] */
    double *pprm[OR_NKIND]; /* [(inst*PK + k)*N + l] */
    double *lim[OR_NKIND];  /* [(l*count + inst)*NU + u] */
};

void or_eval_planes(const struct or_lanes *L, uint32_t l, const double *x, int lim_active, double *g, double *rhs,
                    const or_evctx *ec);
int or_limit_pass(struct or_lanes *L, uint32_t l, const double *x, const double *x_old, int lim_active);
void or_seed_pass(struct or_lanes *L, uint32_t l, const double *x);
/* Prepared constants of instance id of kind on lane l (copied into p). */
void or_load_p(const struct or_lanes *L, int kind, uint32_t id, uint32_t l, double *p);
int or_prepared_count(int kind);
/* Length of the LTE charge tape: one site per capacitor and per diode. */
uint32_t or_tape_len(const struct or_circuit *c);
/* Compiled Verilog-A kinds (va.c). */
int or_va_nport(int kind);
int or_va_nprep(int kind);
int or_va_nsite(int kind);
unsigned or_va_gmask(int kind, int ru);
unsigned or_va_cmask(int kind, int ru);
int or_va_setup(int kind, const double *u, double *p);
int or_va_setup_g(int kind, const double *u, double gmin, double *p);
/* The pattern slot of (row, col), OR_NONE when absent. */
uint32_t or_find_slot(const struct or_circuit *c, uint32_t row, uint32_t col);
/* $simparam("gmin") of the compiled kinds outside the stepping rungs: the
 * deck's gmin (src/core/numerics.zig:386) */
#define OR_SIMPARAM_GMIN 1e-12
int or_va_collapse(int kind, const double *u, const uint32_t *ports, int64_t *rows);
void or_va_eval(int kind, const double *p, const double *x, double *res, double *jac);
void or_va_eval_tran(int kind, const double *p, const double *x, double *res, double *jac, double *q, double *cjac,
                     double *site);
void or_va_eval_charge(int kind, const double *p, const double *x, double *q, double *site);
unsigned or_va_limw(int kind);
int or_va_limit(int kind, const double *p, const double *cur, const double *old, double *lim);
void or_va_seed(int kind, const double *p, const double *cur, double *lim);

double or_fetlim(double vnew, double vold, double vto);
double or_limvds(double vnew, double vold);
double or_pnjlim(double vnew, double vold, double vt, double vcrit, int *icheck);
void or_snap_limited(int nu, unsigned reads, unsigned writes, double *lx, const double *cur);

#endif
