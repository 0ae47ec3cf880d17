/* Device bodies of the oracle. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference compiles every model from Verilog-A with its VerA
 * compiler (OpenFAT/VerA#3fabee3c, vera-0.20.0, not in the reference
 * tree) and evaluates it with forward-mode AD (src/device/eval.zig:107-350,
 * evalRange :780-925). VerA's generated code is absent, so each body here
 * restates its .va source by hand, with the engine's own AD contract:
 *   - a dual number carries the value and d/du for each local unknown;
 *     products, quotients and the transcendental bodies (omath.c) follow
 *     one fixed operation order, which the HIP engine copies;
 *   - `$limit` is ngspice's DEVfetlim / DEVlimvds / DEVpnjlim (devsup.c,
 *     the published SPICE3/ngspice semantics), applied in node space: the
 *     limited image of an instance's unknowns, with the residual corrected
 *     by J (x - x_lim) (eval.zig:825-872) and snapped back within 16 ulps
 *     (eval.zig:1111-1125, snapLimited);
 *   - the MODEINITJCT seed of a cold operating point (mos1.va:293-308)
 *     is the first limited image, so tick 0 linearizes at the seed.
 * Files:
 *   vsource.va  DC and PULSE (waveforms 0 and 1), branch row p,n,br
 *   isource.va  DC
 *   resistor.va, capacitor.va
 *   diode.va    series RS (or collapsed), junction charge; no
 *               breakdown/ikf/ikr/recombination/sidewall
 *   mos1.va     rd = rs = 0 (di/si collapsed onto d/s), static currents,
 *               bulk junctions; Meyer gate charge is not restated yet
 * Parity pinned by: the docs golden DC vector of docs/examples/diode_iv.sp
 * (docs/learn/analyses.md:88-93, i(v1) at 10 and 20 mV to full precision),
 * the FD Jacobian of every body (tests/test_oracle_newton.py, the
 * reference's device/tests/jacobian_fd.zig property) and KCL at every
 * converged point (an independent Shichman-Hodges evaluation, same file). Agreement with VerA's generated mos1 bits: parity
 * unpinned (compiler absent).
 */
#include <math.h>
#include <string.h>

#include "oracle.h"
#include "oracle_internal.h"

/* ------------------------------------------------------------------ dual
 * The reference's Dual (src/device/eval.zig:107-350): the product rule and
 * a quotient's numerator fused (mulAddV is a fused multiply-add on the FMA
 * targets it ships, eval.zig:56-61, 346-348), the quotient's lanes scaled by
 * the reciprocal, guarded sqrt and pow slopes. */

static dual dc(double v) { dual r; memset(&r, 0, sizeof r); r.v = v; return r; }
static dual dvar(double v, int i) { dual r = dc(v); r.d[i] = 1.0; return r; }
static dual dadd(dual a, dual b) { dual r; r.v = a.v + b.v; for (int i = 0; i < DUAL_W; i++) r.d[i] = a.d[i] + b.d[i]; return r; }
static dual dsub(dual a, dual b) { dual r; r.v = a.v - b.v; for (int i = 0; i < DUAL_W; i++) r.d[i] = a.d[i] - b.d[i]; return r; }
static dual dneg(dual a) { dual r; r.v = -a.v; for (int i = 0; i < DUAL_W; i++) r.d[i] = -a.d[i]; return r; }
static dual dmul(dual a, dual b)
{
    dual r;
    r.v = a.v * b.v;
    for (int i = 0; i < DUAL_W; i++) r.d[i] = fma(b.d[i], a.v, a.d[i] * b.v);
    return r;
}
/* the lanes keep the reciprocal (eval.zig:259-268) */
static dual ddiv(dual a, dual b)
{
    dual r;
    const double inv = 1.0 / b.v;
    r.v = a.v / b.v;
    for (int i = 0; i < DUAL_W; i++) r.d[i] = fma(b.d[i], -r.v, a.d[i]) * inv;
    return r;
}
static dual dscale(double s, dual a) { dual r; r.v = s * a.v; for (int i = 0; i < DUAL_W; i++) r.d[i] = s * a.d[i]; return r; }
static dual ddivs(dual a, double s)
{
    dual r;
    const double inv = 1.0 / s;
    r.v = a.v / s;
    for (int i = 0; i < DUAL_W; i++) r.d[i] = a.d[i] * inv;
    return r;
}
static dual daddc(dual a, double c) { a.v = a.v + c; return a; }
static dual dcsub(double c, dual a) { dual r; r.v = c - a.v; for (int i = 0; i < DUAL_W; i++) r.d[i] = -a.d[i]; return r; }
static dual dexp(dual a)
{
    dual r;
    r.v = om_exp(a.v);
    for (int i = 0; i < DUAL_W; i++) r.d[i] = r.v * a.d[i];
    return r;
}
static dual dsqrt(dual a)
{
    dual r;
    r.v = sqrt(a.v);
    double h = r.v > 0.0 ? 0.5 / r.v : 0.0; /* eval.zig:289-294 */
    for (int i = 0; i < DUAL_W; i++) r.d[i] = a.d[i] * h;
    return r;
}
/* x^c for a constant c, slope c x^c / x (c x^(c-1) at x = 0), a non-finite
 * slope becoming 0 (eval.zig:306-316) */
static dual dpowc(dual a, double c)
{
    dual r;
    r.v = om_pow(a.v, c);
    double g = a.v != 0.0 ? c * r.v / a.v : c * om_pow(a.v, c - 1.0);
    if (!isfinite(g)) g = 0.0;
    for (int i = 0; i < DUAL_W; i++) r.d[i] = a.d[i] * g;
    return r;
}
/* Verilog-A min/max on values: the second operand wins only when strictly
 * smaller (larger); the winner keeps its derivatives. */
static dual dminc(dual a, double c) { return c < a.v ? dc(c) : a; }
static dual dmaxc(dual a, double c) { return c > a.v ? dc(c) : a; }

/* ------------------------------------------------------------- limiting */

/* ngspice devsup.c DEVfetlim */
double or_fetlim(double vnew, double vold, double vto)
{
    double vtsthi = fabs(2 * (vold - vto)) + 2;
    double vtstlo = vtsthi / 2 + 2;
    double vtox = vto + 3.5;
    double delv = vnew - vold;
    if (vold >= vto) {
        if (vold >= vtox) {
            if (delv <= 0) {
                if (vnew >= vtox) {
                    if (-delv > vtstlo) vnew = vold - vtstlo;
                } else {
                    vnew = om_max(vnew, vto + 2);
                }
            } else {
                if (delv >= vtsthi) vnew = vold + vtsthi;
            }
        } else {
            if (delv <= 0) vnew = om_max(vnew, vto - 0.5);
            else vnew = om_min(vnew, vto + 4);
        }
    } else {
        if (delv <= 0) {
            if (-delv > vtsthi) vnew = vold - vtsthi;
        } else {
            double vtemp = vto + 0.5;
            if (vnew <= vtemp) {
                if (delv > vtstlo) vnew = vold + vtstlo;
            } else {
                vnew = vtemp;
            }
        }
    }
    return vnew;
}

/* ngspice devsup.c DEVlimvds */
double or_limvds(double vnew, double vold)
{
    if (vold >= 3.5) {
        if (vnew > vold) vnew = om_min(vnew, 3 * vold + 2);
        else if (vnew < 3.5) vnew = om_max(vnew, 2);
    } else {
        if (vnew > vold) vnew = om_min(vnew, 4);
        else vnew = om_max(vnew, -0.5);
    }
    return vnew;
}

/* ngspice devsup.c DEVpnjlim, with the negative-voltage branch */
double or_pnjlim(double vnew, double vold, double vt, double vcrit, int *icheck)
{
    if (vnew > vcrit && fabs(vnew - vold) > vt + vt) {
        if (vold > 0) {
            double arg = 1 + (vnew - vold) / vt;
            if (arg > 0) vnew = vold + vt * om_log(arg);
            else vnew = vcrit;
        } else {
            vnew = vt * om_log(vnew / vt);
        }
        *icheck = 1;
    } else {
        if (vnew < 0) {
            double arg = vold > 0 ? -1 * vold - 1 : 2 * vold - 1;
            if (vnew < arg) {
                vnew = arg;
                *icheck = 1;
            } else {
                *icheck = 0;
            }
        } else {
            *icheck = 0;
        }
    }
    return vnew;
}

/* eval.zig:1116-1125 snapLimited over the unknowns in `reads`/`writes`. */
void or_snap_limited(int nu, unsigned reads, unsigned writes, double *lx, const double *cur)
{
    double scale = 0;
    for (int u = 0; u < nu; u++)
        if (reads >> u & 1) scale = fmax(scale, fabs(cur[u]));
    double tol = 16 * 2.220446049250313e-16 * scale;
    for (int u = 0; u < nu; u++)
        if ((writes & reads) >> u & 1)
            if (fabs(lx[u] - cur[u]) <= tol) lx[u] = cur[u];
}

/* -------------------------------------------------------------- prepare */

static const double KB = 1.38064852e-23, QE = 1.6021766208e-19;
static double given(double v, double dflt) { return v == v ? v : dflt; }

void or_prep_mos1(const double *u, double *p)
{
    const double type = given(u[OR_M1_TYPE], 1.0);
    const double tnom = given(u[OR_M1_TNOM], 27.0);
    const double kp = given(u[OR_M1_KP], 2.0e-5), phi = given(u[OR_M1_PHI], 0.6), gamma = given(u[OR_M1_GAMMA], 0.0);
    const double vto = given(u[OR_M1_VTO], 0.0), tox = given(u[OR_M1_TOX], 1.0e-7), u0 = given(u[OR_M1_U0], 600.0);
    const double nsub = given(u[OR_M1_NSUB], 0.0), nss = given(u[OR_M1_NSS], 0.0), tpg = given(u[OR_M1_TPG], 1.0);
    const double lambda = given(u[OR_M1_LAMBDA], 0.0), ld = given(u[OR_M1_LD], 0.0);
    const double is = given(u[OR_M1_IS], 1.0e-14), js = given(u[OR_M1_JS], 0.0);
    const double cbd = given(u[OR_M1_CBD], 0.0), cbs = given(u[OR_M1_CBS], 0.0), cj = given(u[OR_M1_CJ], 0.0);
    const double cjsw = given(u[OR_M1_CJSW], 0.0), pb = given(u[OR_M1_PB], 0.8), mj = given(u[OR_M1_MJ], 0.5);
    const double mjsw = given(u[OR_M1_MJSW], 0.5), fc = given(u[OR_M1_FC], 0.5);
    const double cgso = given(u[OR_M1_CGSO], 0.0), cgdo = given(u[OR_M1_CGDO], 0.0), cgbo = given(u[OR_M1_CGBO], 0.0);
    const double gmin = given(u[OR_M1_GMIN], 1.0e-12), dtemp = given(u[OR_M1_DTEMP], 0.0);
    const double l = given(u[OR_M1_L], 1.0e-4), w = given(u[OR_M1_W], 1.0e-4);
    const double ad = given(u[OR_M1_AD], 0.0), as = given(u[OR_M1_AS], 0.0), pd = given(u[OR_M1_PD], 0.0), ps = given(u[OR_M1_PS], 0.0);
    const double m = given(u[OR_M1_M], 1.0);
    const int tox_given = u[OR_M1_TOX] == u[OR_M1_TOX], nsub_given = u[OR_M1_NSUB] == u[OR_M1_NSUB];
    const int phi_given = u[OR_M1_PHI] == u[OR_M1_PHI], gamma_given = u[OR_M1_GAMMA] == u[OR_M1_GAMMA];
    const int vto_given = u[OR_M1_VTO] == u[OR_M1_VTO], kp_given = u[OR_M1_KP] == u[OR_M1_KP];
    const int temp_given = u[OR_M1_TEMP] == u[OR_M1_TEMP];

    double tnomk = tnom + 273.15;
    double vtnom = tnomk * (KB / QE);
    double egfet1 = 1.16 - (7.02e-4 * tnomk * tnomk) / (tnomk + 1108.0);
    double coxp = (tox_given && tox > 0.0) ? 3.9 * 8.854214871e-12 / tox : 0.0;
    double phix = phi, gammax = gamma, vtox = vto;
    if (nsub_given && nsub * 1.0e6 > 1.45e16) {
        if (!phi_given) {
            phix = 2.0 * vtnom * om_log(nsub * 1.0e6 / 1.45e16);
            if (phix < 0.1) phix = 0.1;
        }
        if (coxp > 0.0) {
            if (!gamma_given) gammax = sqrt(2.0 * 11.7 * 8.854214871e-12 * 1.6021766208e-19 * nsub * 1.0e6) / coxp;
            if (!vto_given) {
                double fermis = type * 0.5 * phix, wkfng;
                if (tpg != 0.0) {
                    double fermig = tpg * type * 0.5 * egfet1;
                    wkfng = 3.25 + 0.5 * egfet1 - fermig;
                } else {
                    wkfng = 3.2;
                }
                double wkfngs = wkfng - (3.25 + 0.5 * egfet1 + fermis);
                double vfb = wkfngs - nss * 1.0e4 * 1.6021766208e-19 / coxp;
                vtox = vfb + type * (gammax * sqrt(phix) + phix);
            }
        }
    }
    double kpx = (kp_given || coxp <= 0.0) ? kp : u0 * coxp * 1.0e-4;
    double tk = temp_given ? u[OR_M1_TEMP] + 273.15 : OR_TEMPERATURE_K + dtemp;
    double vt = tk * (KB / QE);
    double fact1 = tnomk / 300.15, fact2 = tk / 300.15;
    double egfet = 1.16 - (7.02e-4 * tk * tk) / (tk + 1108.0);
    double arg1 = -egfet1 / (2.0 * KB * tnomk) + 1.1150877 / (2.0 * KB * 300.15);
    double argt = -egfet / (2.0 * KB * tk) + 1.1150877 / (2.0 * KB * 300.15);
    double pbfact1 = -2.0 * vtnom * (1.5 * om_log(fact1) + QE * arg1);
    double pbfact = -2.0 * vt * (1.5 * om_log(fact2) + QE * argt);
    double ratio4 = (tk / tnomk) * sqrt(tk / tnomk);
    double kpt = kpx / ratio4;
    double phio = (phix - pbfact1) / fact1;
    double phit = fact2 * phio + pbfact;
    double vbit = vtox - type * (gammax * sqrt(phix)) + 0.5 * (egfet1 - egfet) + type * 0.5 * (phit - phix);
    double ist = is * om_exp(-egfet / vt + egfet1 / vtnom);
    double jst = js * om_exp(-egfet / vt + egfet1 / vtnom);
    double pbo = (pb - pbfact1) / fact1;
    double pbt = fact2 * pbo + pbfact;
    double gmaold = (pb - pbo) / pbo, gmanew = (pbt - pbo) / pbo;
    double cfb = (1.0 + mj * (4.0e-4 * (tk - 300.15) - gmanew)) / (1.0 + mj * (4.0e-4 * (tnomk - 300.15) - gmaold));
    double cfsw = (1.0 + mjsw * (4.0e-4 * (tk - 300.15) - gmanew)) / (1.0 + mjsw * (4.0e-4 * (tnomk - 300.15) - gmaold));
    double leff = l - 2.0 * ld;
    if (leff < 1.0e-10) leff = 1.0e-10;
    double beta = kpt * m * w / leff;
    double cox = coxp * w * leff * m;
    double isatd, isats;
    if (js > 0.0 && ad > 0.0 && as > 0.0) {
        isatd = jst * ad * m;
        isats = jst * as * m;
    } else {
        isatd = ist * m;
        isats = ist * m;
    }
    double czbd = cbd > 0.0 ? cbd * cfb * m : cj * cfb * ad * m;
    double czbs = cbs > 0.0 ? cbs * cfb * m : cj * cfb * as * m;
    double czbdsw = cjsw * cfsw * pd * m, czbssw = cjsw * cfsw * ps * m;
    double depcap = fc * pbt, argf = 1.0 - fc;
    double sargb = om_pow(argf, -mj), sargs = om_pow(argf, -mjsw);
    double f2d = czbd * (1.0 - fc * (1.0 + mj)) * sargb / argf + czbdsw * (1.0 - fc * (1.0 + mjsw)) * sargs / argf;
    double f3d = czbd * mj * sargb / (argf * pbt) + czbdsw * mjsw * sargs / (argf * pbt);
    double f4d = czbd * pbt * (1.0 - argf * sargb) / (1.0 - mj) + czbdsw * pbt * (1.0 - argf * sargs) / (1.0 - mjsw)
                 - 0.5 * f3d * depcap * depcap - depcap * f2d;
    double f2s = czbs * (1.0 - fc * (1.0 + mj)) * sargb / argf + czbssw * (1.0 - fc * (1.0 + mjsw)) * sargs / argf;
    double f3s = czbs * mj * sargb / (argf * pbt) + czbssw * mjsw * sargs / (argf * pbt);
    double f4s = czbs * pbt * (1.0 - argf * sargb) / (1.0 - mj) + czbssw * pbt * (1.0 - argf * sargs) / (1.0 - mjsw)
                 - 0.5 * f3s * depcap * depcap - depcap * f2s;
    p[M1P_TYPE] = type;
    p[M1P_VT] = vt;
    p[M1P_VBIT] = vbit;
    p[M1P_PHIT] = phit;
    p[M1P_GAMMA] = gammax;
    p[M1P_BETA] = beta;
    p[M1P_LAMBDA] = lambda;
    p[M1P_ISATS] = isats;
    p[M1P_ISATD] = isatd;
    p[M1P_GMIN] = gmin;
    p[M1P_VCRITS] = vt * om_log(vt / (1.4142135623730951 * isats));
    p[M1P_VCRITD] = vt * om_log(vt / (1.4142135623730951 * isatd));
    p[M1P_COX] = cox;
    p[M1P_CZBD] = czbd;
    p[M1P_CZBS] = czbs;
    p[M1P_CZBDSW] = czbdsw;
    p[M1P_CZBSSW] = czbssw;
    p[M1P_PBT] = pbt;
    p[M1P_MJ] = mj;
    p[M1P_MJSW] = mjsw;
    p[M1P_FC] = fc;
    p[M1P_F2D] = f2d;
    p[M1P_F3D] = f3d;
    p[M1P_F4D] = f4d;
    p[M1P_F2S] = f2s;
    p[M1P_F3S] = f3s;
    p[M1P_F4S] = f4s;
    p[M1P_CGSO] = cgso * w * m;
    p[M1P_CGDO] = cgdo * w * m;
    p[M1P_CGBO] = cgbo * leff * m;
    p[M1P_SEED] = type * vbit + gammax * sqrt(phit);
}

void or_prep_diode(const double *u, double *p)
{
    const double is = given(u[OR_DIO_IS], 1e-14), n = given(u[OR_DIO_N], 1.0), cjo = given(u[OR_DIO_CJO], 0.0);
    const double vj = given(u[OR_DIO_VJ], 1.0), m = given(u[OR_DIO_M], 0.5), fc = given(u[OR_DIO_FC], 0.5);
    const double tt = given(u[OR_DIO_TT], 0.0), gmin = given(u[OR_DIO_GMIN], 1e-12), area = given(u[OR_DIO_AREA], 1.0);
    const double tnom = given(u[OR_DIO_TNOM], 27.0), dtemp = given(u[OR_DIO_DTEMP], 0.0), xti = given(u[OR_DIO_XTI], 3.0);
    const int temp_given = u[OR_DIO_TEMP] == u[OR_DIO_TEMP];
    double tk = temp_given ? u[OR_DIO_TEMP] + 273.15 : OR_TEMPERATURE_K + dtemp;
    double tnomk = tnom + 273.15;
    double ratio = tk / tnomk, dtk = tk - tnomk;
    double vt = KB / QE * tk, vtnom = KB / QE * tnomk;
    double vte = n * vt;
    double ae = area;
    double egp = given(u[OR_DIO_EG], 1.11);
    double egfet = 1.16 - (7.02e-4 * tk * tk) / (tk + 1108.0);
    double egfet1 = 1.16 - (7.02e-4 * tnomk * tnomk) / (tnomk + 1108.0);
    const double tm1 = given(u[OR_DIO_TM1], 0.0), tm2 = given(u[OR_DIO_TM2], 0.0);
    const double rs = given(u[OR_DIO_RS], 0.0), trs = given(u[OR_DIO_TRS], 0.0), trs2 = given(u[OR_DIO_TRS2], 0.0);
    double mt = m * (1.0 + tm1 * dtk + tm2 * dtk * dtk);
    double fact1 = tnomk / 300.15, fact2 = tk / 300.15;
    double arg1 = -egfet1 / (2.0 * KB * tnomk) + 1.1150877 / (2.0 * KB * 300.15);
    double argt = -egfet / (2.0 * KB * tk) + 1.1150877 / (2.0 * KB * 300.15);
    double pbfact1 = -2.0 * vtnom * (1.5 * om_log(fact1) + QE * arg1);
    double pbfact = -2.0 * vt * (1.5 * om_log(fact2) + QE * argt);
    double pbo = (vj - pbfact1) / fact1;
    double gmaold = (vj - pbo) / pbo;
    double vj_t = pbfact + fact2 * pbo;
    double gmanew = (vj_t - pbo) / pbo;
    double cjo_t = cjo * (1.0 + mt * (400e-6 * (tk - 300.15) - gmanew)) / (1.0 + mt * (400e-6 * (tnomk - 300.15) - gmaold));
    cjo_t = cjo_t * ae;
    double is_t = is * ae * om_exp((ratio - 1.0) * egp / vte + (xti / n) * om_log(ratio));
    double issw_t = 0.0 * 0.0 * om_exp((ratio - 1.0) * egp / vt + (xti / 1.0) * om_log(ratio));
    is_t = is_t + issw_t;
    p[DIOP_IS] = is_t;
    p[DIOP_VTE] = vte;
    p[DIOP_GMIN] = gmin;
    p[DIOP_VCRIT] = vte * om_log(vte / (1.4142135623730951 * (is_t + 0.0)));
    p[DIOP_CJO] = cjo_t;
    p[DIOP_VJ] = vj_t;
    p[DIOP_MT] = mt;
    p[DIOP_FC] = fc;
    p[DIOP_TT] = tt;
    p[DIOP_QON] = cjo > 0.0;
    p[DIOP_RSON] = rs > 0.0;
    /* diode.va:259-262, RS(T) floored at 1 nOhm */
    double rs_t = rs * (1.0 + trs * dtk + trs2 * dtk * dtk) / ae;
    p[DIOP_RST] = rs_t < 1.0e-9 ? 1.0e-9 : rs_t;
    (void)dtemp;
}

/* ----------------------------------------------------------------- eval */

/* mos1.va:106-116 */
static dual jcur(dual v, double isat, double vtx, double gm)
{
    if (v.v > -3.0 * vtx) {
        dual w = ddivs(v, vtx);
        dual u = dminc(w, 80.0);
        dual e = dexp(u);
        dual t = dmul(e, dadd(dc(1.0), dsub(w, u)));
        return dadd(dscale(isat, daddc(t, -1.0)), dscale(gm, v));
    }
    return daddc(dscale(gm, v), -isat);
}

/* mos1.va:120-127 */
static dual sbody(dual vbx, double phit)
{
    if (vbx.v <= 0.0) return dsqrt(dcsub(phit, vbx));
    double sp = sqrt(phit);
    return dmaxc(dcsub(sp, ddivs(vbx, 2.0 * sp)), 0.0);
}

/* Static MOS1 evaluation at local voltages lx (d, g, s, b): residual rows
 * and dF/du (mos1.va:299-345, 447-449, with collapsed di/si). */
void or_eval_mos1(const double *p, const double *lx, or_devout *o)
{
    const double type = p[M1P_TYPE];
    dual d = dvar(lx[0], 0), g = dvar(lx[1], 1), s = dvar(lx[2], 2), b = dvar(lx[3], 3);
    dual vgs = dscale(type, dsub(g, s));
    dual vds = dscale(type, dsub(d, s));
    dual vbs = dscale(type, dsub(b, s));
    dual vbd = dscale(type, dsub(b, d));
    dual vgd = dsub(vgs, vds);
    dual ibs = jcur(vbs, p[M1P_ISATS], p[M1P_VT], p[M1P_GMIN]);
    dual ibd = jcur(vbd, p[M1P_ISATD], p[M1P_VT], p[M1P_GMIN]);
    double mode;
    dual vbx, vgsx;
    if (vds.v >= 0.0) {
        mode = 1.0;
        vbx = vbs;
        vgsx = vgs;
    } else {
        mode = -1.0;
        vbx = vbd;
        vgsx = vgd;
    }
    dual vdsx = dscale(mode, vds);
    dual sarg = sbody(vbx, p[M1P_PHIT]);
    if (vbx.v > 0.0 && sarg.v > 0.0) {
        double sarg0 = sarg.v, vbx0 = vbx.v;
        sarg = dcsub(sarg0, ddivs(daddc(vbx, -vbx0), 2.0 * sarg0));
    }
    dual von = dadd(dc(p[M1P_VBIT] * type), dscale(p[M1P_GAMMA], sarg));
    dual vgst = dsub(vgsx, von);
    dual bp = dscale(p[M1P_BETA], dadd(dc(1.0), dscale(p[M1P_LAMBDA], vdsx)));
    dual id;
    if (vgst.v <= 0.0) id = dc(0.0);
    else if (vgst.v <= vdsx.v) id = dmul(dmul(dscale(0.5, bp), vgst), vgst);
    else id = dmul(dmul(bp, vdsx), dsub(vgst, dscale(0.5, vdsx)));
    dual cd = dscale(type * mode, id);
    dual tibd = dscale(type, ibd), tibs = dscale(type, ibs);
    dual rd = dsub(cd, tibd);
    dual rs = dsub(dneg(cd), tibs);
    dual rb = dadd(tibd, tibs);
    o->res[0] = rd.v;
    o->res[1] = 0.0;
    o->res[2] = rs.v;
    o->res[3] = rb.v;
    for (int c = 0; c < 4; c++) {
        o->jac[0][c] = rd.d[c];
        o->jac[1][c] = 0.0;
        o->jac[2][c] = rs.d[c];
        o->jac[3][c] = rb.d[c];
    }
}

/* MOS1 limit in node space (mos1.va:286-308 with ngspice mos1load.c's
 * order): returns 1 when pnjlim clamped (the iterate may not converge). */
int or_limit_mos1(const double *p, const double *cur, const double *old, double *lim)
{
    const double type = p[M1P_TYPE];
    double vgs_o = type * (old[1] - old[2]), vds_o = type * (old[0] - old[2]);
    double vbs_o = type * (old[3] - old[2]), vbd_o = type * (old[3] - old[0]);
    double vgd_o = vgs_o - vds_o;
    double vgs = type * (cur[1] - cur[2]), vds = type * (cur[0] - cur[2]);
    double vbs = type * (cur[3] - cur[2]), vbd = type * (cur[3] - cur[0]);
    double vgd = vgs - vds;
    double vbx_o = vds_o >= 0.0 ? vbs_o : vbd_o;
    double sb;
    if (vbx_o <= 0.0) sb = sqrt(p[M1P_PHIT] - vbx_o);
    else sb = om_max(sqrt(p[M1P_PHIT]) - vbx_o / (2.0 * sqrt(p[M1P_PHIT])), 0.0);
    double von = p[M1P_VBIT] * type + p[M1P_GAMMA] * sb;
    if (vds_o >= 0.0) {
        vgs = or_fetlim(vgs, vgs_o, von);
        vds = vgs - vgd;
        vds = or_limvds(vds, vds_o);
    } else {
        vgd = or_fetlim(vgd, vgd_o, von);
        vds = vgs - vgd;
        vds = -or_limvds(-vds, -vds_o);
        vgs = vgd + vds;
    }
    int check = 0;
    if (vds >= 0.0) {
        vbs = or_pnjlim(vbs, vbs_o, p[M1P_VT], p[M1P_VCRITS], &check);
    } else {
        vbd = or_pnjlim(vbd, vbd_o, p[M1P_VT], p[M1P_VCRITD], &check);
        vbs = vbd + vds;
    }
    double s = cur[2];
    lim[0] = s + type * vds;
    lim[1] = s + type * vgs;
    lim[2] = s;
    lim[3] = s + type * vbs;
    or_snap_limited(4, 0xf, 0xb, lim, cur);
    return check;
}

/* MODEINITJCT image: vgs = type*vbit + gamma*sqrt(phit), vds = 0, vbs = -1
 * in the type frame, anchored at the source node. */
void or_seed_mos1(const double *p, const double *cur, double *lim)
{
    const double type = p[M1P_TYPE];
    double s = cur[2];
    lim[0] = s;
    lim[1] = s + type * p[M1P_SEED];
    lim[2] = s;
    lim[3] = s + type * -1.0;
}

/* diode.va:282-372 without ikf/ikr/isr/bv/sidewall. Unknowns a, c, ai (ai
 * is a's row when RS is 0). Junction across (ai, c) at the limited vd:
 * I = ihi + gmin vd into ai, out of c; with RS on, I(a, ai) = V(a, ai) / RS(T).
 * tran: the junction charge qj (depletion with the fc continuation, plus
 * tt * I) on rows ai (+) and c (-). */
void or_eval_diode(const double *p, const double *lx, or_devout *o, int tran)
{
    dual a = dvar(lx[0], 0), c = dvar(lx[1], 1), ai = dvar(lx[2], 2);
    dual vd = dsub(ai, c);
    const double vte = p[DIOP_VTE], is_t = p[DIOP_IS];
    dual id;
    if (vd.v >= -3.0 * vte) {
        dual u = ddivs(vd, vte);
        dual ulim = dminc(u, 80.0);
        id = dscale(is_t, daddc(dmul(dexp(ulim), dadd(dc(1.0), dsub(u, ulim))), -1.0));
    } else {
        dual t = dpowc(ddiv(dc(3.0 * vte), dscale(2.718281828459045, vd)), 3.0);
        id = dscale(-is_t, dadd(dc(1.0), t));
    }
    dual ij = dadd(id, dscale(p[DIOP_GMIN], vd));
    dual ir = dc(0.0);
    if (p[DIOP_RSON] != 0.0) ir = ddivs(dsub(a, ai), p[DIOP_RST]);
    dual rai = dsub(ij, ir);
    o->res[0] = ir.v;
    o->res[1] = -ij.v;
    o->res[2] = rai.v;
    for (int k = 0; k < 3; k++) {
        o->jac[0][k] = ir.d[k];
        o->jac[1][k] = -ij.d[k];
        o->jac[2][k] = rai.d[k];
    }
    if (!tran) return;
    dual qj = dc(0.0);
    if (p[DIOP_QON] != 0.0) {
        const double cj = p[DIOP_CJO], vj = p[DIOP_VJ], mt = p[DIOP_MT], fc = p[DIOP_FC];
        const double vfc = fc * vj;
        if (vd.v < vfc) {
            /* cjo(T) vj(T) (1 - (1 - vd/vj)^(1-M)) / (1 - M) */
            dual base = dcsub(1.0, ddivs(vd, vj));
            qj = ddivs(dscale(cj * vj, dcsub(1.0, dpowc(base, 1.0 - mt))), 1.0 - mt);
        } else {
            double f1 = vj * (1.0 - om_pow(1.0 - fc, 1.0 - mt)) / (1.0 - mt);
            double f2 = om_pow(1.0 - fc, 1.0 + mt);
            double f3 = 1.0 - fc * (1.0 + mt);
            dual t1 = dscale(f3, daddc(vd, -vfc));
            dual t2 = dscale(mt / (2.0 * vj), daddc(dmul(vd, vd), -(vfc * vfc)));
            qj = dscale(cj, dadd(dc(f1), dscale(1.0 / f2, dadd(t1, t2))));
        }
    }
    qj = dadd(qj, dscale(0.0, vd)); /* ccox vd, ccox = 0 below level 3 */
    if (p[DIOP_TT] > 0.0) qj = dadd(qj, dscale(p[DIOP_TT], ij));
    o->q[1] = -qj.v;
    o->q[2] = qj.v;
    for (int k = 0; k < 3; k++) {
        o->cjac[1][k] = -qj.d[k];
        o->cjac[2][k] = qj.d[k];
    }
}

/* $limit(V(ai, c), "pnjlim"): reads c and ai, writes ai. */
int or_limit_diode(const double *p, const double *cur, const double *old, double *lim)
{
    int check = 0;
    double vd = or_pnjlim(cur[2] - cur[1], old[2] - old[1], p[DIOP_VTE], p[DIOP_VCRIT], &check);
    lim[0] = cur[0];
    lim[1] = cur[1];
    lim[2] = cur[1] + vd;
    or_snap_limited(3, 0x6, 0x4, lim, cur);
    return check;
}

void or_seed_diode(const double *p, const double *cur, double *lim)
{
    lim[0] = cur[0];
    lim[1] = cur[1];
    lim[2] = cur[1] + p[DIOP_VCRIT];
}

/* ---------------------------------------------------------------- sources */

static double vmax(double a, double b) { return a > b ? a : b; }

/* The PULSE widths of models/vsource.va:44-50 (P_TR, P_TF, P_PW, P_PER). */
static void pulse_widths(const double *p, double *tr, double *tf, double *pw, double *per)
{
    *tr = vmax(p[SRCP_TR], 1.0e-12);
    *tf = vmax(p[SRCP_TF], 1.0e-12);
    *pw = vmax(p[SRCP_PW], 0.0) + vmax(-p[SRCP_PW], 0.0) * 1.0e30;
    *per = vmax(p[SRCP_PER], *tr + *pw + *tf + 1.0e-12);
}

/* models/vsource.va:260-299, PWL: te = t - td folded into [r, t_last] by
 * r=, v_first before the first point, v_last after the last, else the
 * segment bisection finds, its ramp at most 1 fs steep. */
static double pwl_at(const double *p, double t)
{
    const int n = (int)p[SRCP_PWL_LEN];
    if (n <= 0) return p[SRCP_DC];
    const double *T = p + SRCP_PWL_T, *V = p + SRCP_PWL_V;
    double te = t - p[SRCP_PWL_TD];
    if (p[SRCP_PWL_R] >= 0.0 && n > 1 && te > T[n - 1] && T[n - 1] > p[SRCP_PWL_R]) {
        const double span = T[n - 1] - p[SRCP_PWL_R];
        if (span > 0.0) {
            const double excess = te - p[SRCP_PWL_R];
            te = p[SRCP_PWL_R] + (excess - floor(excess / span) * span);
        }
    }
    if (te < 0.0 || te <= T[0]) return V[0];
    if (te >= T[n - 1]) return V[n - 1];
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (T[mid] > te) hi = mid;
        else lo = mid;
    }
    return V[lo] + (V[hi] - V[lo]) * (te - T[lo]) / vmax(T[hi] - T[lo], 1.0e-15);
}

/* models/vsource.va:226-299: DC, or PULSE at t (PHASE 0), SIN, PWL. */
double or_source_value(const double *p, double t)
{
    if (p[SRCP_WAVE] == 4.0) return pwl_at(p, t);
    if (p[SRCP_WAVE] == 5.0) {
        /* SFFM, models/vsource.va:304-317 */
        double mdi = p[SRCP_FMDI];
        if (mdi > p[SRCP_FFC] / p[SRCP_FFM]) mdi = p[SRCP_FFC] / p[SRCP_FFM];
        else if (mdi < 0.0) mdi = 0.0;
        const double tf = t - p[SRCP_FTD];
        if (tf <= 0.0) return 0.0;
        return p[SRCP_FVO] + p[SRCP_FVA] * om_sin((VA_TWO_PI * p[SRCP_FFC] * tf + p[SRCP_FPHC] * VA_PI / 180.0) +
                                                  mdi * om_sin(VA_TWO_PI * p[SRCP_FFM] * tf + p[SRCP_FPHM] * VA_PI / 180.0));
    }
    if (p[SRCP_WAVE] == 6.0) {
        /* AM, vsource.va and isource.va alike */
        if (!(t >= p[SRCP_ATD])) return 0.0;
        const double env = p[SRCP_AVO] + om_sin(VA_TWO_PI * p[SRCP_AMF] * t + p[SRCP_APHS] * VA_PI / 180.0);
        const double car = om_sin(VA_TWO_PI * p[SRCP_AFC] * t + p[SRCP_APHC] * VA_PI / 180.0);
        return p[SRCP_AVA] * env * car;
    }
    if (p[SRCP_WAVE] == 3.0) {
        /* EXP, models/vsource.va:252-258 */
        double out = p[SRCP_EV1];
        const double e1 = t - p[SRCP_ETD1], e2 = t - p[SRCP_ETD2];
        if (e1 > 0.0)
            out = out + (p[SRCP_EV2] - p[SRCP_EV1]) * (1.0 - om_exp(om_min(-e1 / vmax(p[SRCP_ETAU1], 1.0e-15), 80.0)));
        if (e2 > 0.0)
            out = out + (p[SRCP_EV1] - p[SRCP_EV2]) * (1.0 - om_exp(om_min(-e2 / vmax(p[SRCP_ETAU2], 1.0e-15), 80.0)));
        return out;
    }
    if (p[SRCP_WAVE] == 2.0) {
        /* SIN, models/vsource.va:233-244 */
        const double t_sin = t - p[SRCP_STD];
        if (t_sin < 0.0) return p[SRCP_SVO] + p[SRCP_SVA] * om_sin(p[SRCP_SPHASE] * VA_PI / 180.0);
        return p[SRCP_SVO] + p[SRCP_SVA] * om_sin(VA_TWO_PI * p[SRCP_SFREQ] * t_sin + p[SRCP_SPHASE] * VA_PI / 180.0) *
                                 om_exp(om_min(-p[SRCP_STHETA] * t_sin, 80.0));
    }
    if (p[SRCP_WAVE] != 1.0) return p[SRCP_DC];
    double tr, tf, pw, per;
    pulse_widths(p, &tr, &tf, &pw, &per);
    const double v1 = p[SRCP_V1], v2 = p[SRCP_V2];
    const double t_rel = t - p[SRCP_TD];
    double t_mod = t_rel + (0.0 / 360.0) * per;
    t_mod = t_mod - floor(t_mod / per) * per;
    t_mod = vmax(t_mod, 0.0);
    if (t_rel < 0.0) return v1;
    if (t_mod < tr) return v1 + (v2 - v1) * (t_mod / tr);
    if (t_mod < tr + pw) return v2;
    if (t_mod < tr + pw + tf) return v2 - (v2 - v1) * ((t_mod - (tr + pw)) / tf);
    return v1;
}

/* src/device/eval.zig:1636-1650 spicePulseBreak: the corner after tq,
 * rounded from the accepted time t_acc; inf when none (not PULSE, or the
 * rounded time is not after tq). */
/* The fire time after t of `@(timer(start, period))` (LRM 5.10.5): start
 * + k period for the first k that lands strictly after t. VerA generates
 * vsource's nextBreakpoint from the four PULSE corner timers of
 * models/vsource.va:137-140; its rounding of start + k period is not in the
 * reference tree (parity pinned through the point counts of the decks
 * whose second period needs it, docs/learn/models.md:95-102). */
static double timer_next(double start, double period, double t)
{
    if (start > t) return start;
    const double k = floor((t - start) / period) + 1.0;
    double next = start + k * period;
    if (!(next > t)) next = start + (k + 1.0) * period;
    return next;
}

/* vsource's nextBreakpoint: the earliest of the four PULSE corner timers
 * after t, whatever the waveform (a non-PULSE source parks TD at 1e30). */
double or_pulse_timer_break(const double *p, double t)
{
    double tr, tf, pw, per;
    pulse_widths(p, &tr, &tf, &pw, &per);
    const double td = p[SRCP_TD];
    double best = timer_next(td, per, t);
    best = fmin(best, timer_next(td + tr, per, t));
    best = fmin(best, timer_next(td + tr + pw, per, t));
    best = fmin(best, timer_next(td + tr + pw + tf, per, t));
    return best;
}

double or_pulse_break(const double *p, double t_acc, double tq)
{
    if (p[SRCP_WAVE] != 1.0) return INFINITY;
    double tr, tf, pw, per;
    pulse_widths(p, &tr, &tf, &pw, &per);
    double time = t_acc - p[SRCP_TD];
    if (time >= per) time -= per * floor(time / per);
    const double atime = time + (tq - t_acc);
    const double corner = atime < 0.0 ? 0.0 : atime < tr ? tr : atime < tr + pw ? tr + pw : atime < tr + pw + tf ? tr + pw + tf : per;
    const double bp = t_acc + (corner - time);
    return bp > tq ? bp : INFINITY;
}

/* src/device/eval.zig:1661-1697 spicePwlBreak: 1 with *bp set (a corner, or
 * INFINITY past a table that does not repeat), 0 for none (the timers then
 * answer). */
static int pwl_break(const double *p, double t_acc, double tq, double *bp)
{
    const int n = (int)p[SRCP_PWL_LEN];
    const double *T = p + SRCP_PWL_T, r = p[SRCP_PWL_R];
    const int repeats = r >= 0.0 && n > 1 && T[n - 1] > r;
    const double period = T[n - 1] - r;
    double time = t_acc - p[SRCP_PWL_TD];
    if (time > T[n - 1]) {
        if (!repeats) {
            *bp = INFINITY;
            return 1;
        }
        time = time - r;
        time = time - period * floor(time / period);
        time = time + r;
    }
    const double atime = time + (tq - t_acc);
    int k = 0;
    while (k < n && !(T[k] > atime)) k++;
    double corner;
    if (k < n) {
        corner = T[k];
    } else {
        if (!repeats) {
            *bp = INFINITY;
            return 1;
        }
        k = 0;
        while (k < n && !(T[k] + period > atime)) k++;
        if (k == n) return 0;
        corner = T[k] + period;
    }
    const double b = t_acc + (corner - time);
    if (!(b > tq)) return 0;
    *bp = b;
    return 1;
}

/* srcBreak (src/device/eval.zig:1592-1598): the PULSE corner, else the PWL
 * corner, else the source's timers: the four PULSE corners and one
 * one-shot timer per PWL slot at t_k + td (models/vsource.va:137-204). */
double or_src_break(const double *p, double t_acc, double tq)
{
    const double pb = or_pulse_break(p, t_acc, tq);
    if (pb != INFINITY) return pb;
    double b;
    if (p[SRCP_WAVE] == 4.0 && p[SRCP_PWL_LEN] > 0 && pwl_break(p, t_acc, tq, &b)) return b;
    double best = or_pulse_timer_break(p, tq);
    for (int k = 0; k < OR_PWL_CAP; k++) {
        const double at = p[SRCP_PWL_T + k] + p[SRCP_PWL_TD];
        if (at > tq && at < best) best = at;
    }
    return best;
}
