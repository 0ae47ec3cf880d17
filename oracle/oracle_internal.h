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
enum { DIOP_IS = 0, DIOP_VTE, DIOP_GMIN, DIOP_VCRIT, DIOP_CJO, DIOP_VJ, DIOP_MT, DIOP_FC, DIOP_TT, DIOP_K };

void or_prep_mos1(const double *u, double *p);
void or_prep_diode(const double *u, double *p);
void or_eval_mos1(const double *p, const double *lx, or_devout *o);
void or_eval_diode(const double *p, const double *lx, or_devout *o);
int or_limit_mos1(const double *p, const double *cur, const double *old, double *lim);
int or_limit_diode(const double *p, const double *cur, const double *old, double *lim);
void or_seed_mos1(const double *p, const double *cur, double *lim);
void or_seed_diode(const double *p, const double *cur, double *lim);
double or_fetlim(double vnew, double vold, double vto);
double or_limvds(double vnew, double vold);
double or_pnjlim(double vnew, double vold, double vt, double vcrit, int *icheck);

#endif
