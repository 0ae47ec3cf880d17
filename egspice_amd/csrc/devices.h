// Device models of the lane engine: parameter tables, temperature and
// geometry preparation, static evaluation with forward-mode AD, and SPICE
// limiting. One body serves the host tools and the gfx950 kernels.
//
// The reference compiles each model from Verilog-A with VerA (absent from
// its tree); these bodies restate the .va sources by hand:
//   models/vsource.va  (DC; branch unknown p, n, br)      vsource.va:230
//   models/isource.va  (DC)
//   models/resistor.va, models/capacitor.va
//   models/diode.va    (rs = 0, the ai node collapsed)     diode.va:190-236
//   models/mos1.va     (rd = rs = 0, di/si collapsed)      mos1.va:151-449
// Limiting is ngspice's DEVfetlim / DEVlimvds / DEVpnjlim (devsup.c) in
// node space: the limited image of an instance's unknowns, the residual
// corrected by J (x - x_lim) (src/device/eval.zig:825-872) and snapped
// back within 16 ulps (eval.zig:1111-1125).
#pragma once

#include "dual.h"
#include "va_models.h"

namespace esp {

// K_PSP103 and the kinds after it are compiled from Verilog-A (va_models.h).
enum Kind : int {
    K_VSRC = 0, K_ISRC = 1, K_RES = 2, K_CAP = 3, K_DIODE = 4, K_MOS1 = 5, K_PSP103 = 6, K_BSIM4 = 7, K_NKIND = 8
};
constexpr bool kind_is_va(int kind) { return kind >= K_PSP103; }
// The kernels are instantiated per set of compiled kinds a template holds
// (a bit per kind), so that a template's kernel carries only its own
// models' evaluations and register budget.
enum { VA_PSP = 1, VA_B4 = 2, VA_SETS = 4 };
constexpr int va_bit(int kind) { return kind == K_PSP103 ? VA_PSP : kind == K_BSIM4 ? VA_B4 : 0; }

// User parameter slots (NaN = not given).
// waveform: 0 DC, 1 PULSE(V1 V2 TD TR TF PW PER), 2 SIN(VO VA FREQ TD THETA
// PHASE), 3 EXP(V1 V2 TD1 TAU1 TD2 TAU2), 4 PWL(T1 V1 ...) with its point
// count, r= and td=; the table holds
// PWL_CAP points (vsource.va's holds 64; a longer deck table is refused).
constexpr int PWL_CAP = 16;
enum { VSRC_DC = 0, VSRC_WAVE, VSRC_V1, VSRC_V2, VSRC_TD, VSRC_TR, VSRC_TF, VSRC_PW, VSRC_PER, VSRC_SVO, VSRC_SVA,
       VSRC_SFREQ, VSRC_STD, VSRC_STHETA, VSRC_SPHASE, VSRC_PWL_LEN, VSRC_PWL_R, VSRC_PWL_TD, VSRC_PWL_T,
       VSRC_PWL_V = VSRC_PWL_T + PWL_CAP, VSRC_EV1 = VSRC_PWL_V + PWL_CAP, VSRC_EV2, VSRC_ETD1, VSRC_ETAU1, VSRC_ETD2,
       VSRC_ETAU2, VSRC_FVO, VSRC_FVA, VSRC_FFM, VSRC_FMDI, VSRC_FFC, VSRC_FTD, VSRC_FPHM, VSRC_FPHC, VSRC_AVA, VSRC_AVO,
       VSRC_AMF, VSRC_AFC, VSRC_ATD, VSRC_APHC, VSRC_APHS, VSRC_K };
static_assert(VSRC_EV1 == 50 && VSRC_FVO == 56 && VSRC_AVA == 64 && VSRC_K == 71,
              "the deck reader writes EXP, SFFM and AM arguments into these slots");
static_assert(VSRC_PWL_LEN == 15 && VSRC_PWL_R == 16 && VSRC_PWL_TD == 17 && VSRC_PWL_T == 18 && VSRC_PWL_V == 34,
              "the deck reader writes PWL tables into these slots");
// Prepared source constants: the user slots with vsource.va's defaults
// (models/vsource.va:58-66: 0, and -1 sentinels for TR, TF, PW, PER).
enum { SRCP_DC = 0, SRCP_WAVE, SRCP_V1, SRCP_V2, SRCP_TD, SRCP_TR, SRCP_TF, SRCP_PW, SRCP_PER, SRCP_SVO, SRCP_SVA,
       SRCP_SFREQ, SRCP_STD, SRCP_STHETA, SRCP_SPHASE, SRCP_PWL_LEN, SRCP_PWL_R, SRCP_PWL_TD, SRCP_PWL_T,
       SRCP_PWL_V = SRCP_PWL_T + PWL_CAP, SRCP_EV1 = SRCP_PWL_V + PWL_CAP, SRCP_EV2, SRCP_ETD1, SRCP_ETAU1, SRCP_ETD2,
       SRCP_ETAU2, SRCP_FVO, SRCP_FVA, SRCP_FFM, SRCP_FMDI, SRCP_FFC, SRCP_FTD, SRCP_FPHM, SRCP_FPHC, SRCP_AVA, SRCP_AVO,
       SRCP_AMF, SRCP_AFC, SRCP_ATD, SRCP_APHC, SRCP_APHS, SRCP_K };
constexpr double VA_PI = 3.14159265358979323846, VA_TWO_PI = 6.28318530717958647652;  // constants.vams
// the current source's parameters are the voltage source's, slot for slot
// (models/isource.va:63-130 keeps the waveform names' `v` spelling)
enum { ISRC_DC = 0, ISRC_K = VSRC_K };
enum { RES_R = 0, RES_K = 1 };
enum { CAP_C = 0, CAP_K = 1 };
enum { DIO_IS = 0, DIO_N, DIO_CJO, DIO_VJ, DIO_M, DIO_FC, DIO_TT, DIO_GMIN, DIO_AREA, DIO_TNOM, DIO_TEMP, DIO_DTEMP,
       DIO_EG, DIO_XTI, DIO_RS, DIO_TRS, DIO_TRS2, DIO_TM1, DIO_TM2, DIO_K };
enum { M1_TYPE = 0, M1_VTO, M1_KP, M1_GAMMA, M1_PHI, M1_LAMBDA, M1_TOX, M1_LD, M1_U0, M1_NSUB, M1_NSS, M1_TPG, M1_IS,
       M1_JS, M1_CBD, M1_CBS, M1_CJ, M1_CJSW, M1_PB, M1_MJ, M1_MJSW, M1_FC, M1_CGSO, M1_CGDO, M1_CGBO, M1_TNOM, M1_GMIN,
       M1_TEMP, M1_DTEMP, M1_L, M1_W, M1_AD, M1_AS, M1_PD, M1_PS, M1_M, M1_K };

// Prepared constants per instance and lane.
enum { M1P_TYPE = 0, M1P_VT, M1P_VBIT, M1P_PHIT, M1P_GAMMA, M1P_BETA, M1P_LAMBDA, M1P_ISATS, M1P_ISATD, M1P_GMIN,
       M1P_VCRITS, M1P_VCRITD, M1P_COX, M1P_CZBD, M1P_CZBS, M1P_CZBDSW, M1P_CZBSSW, M1P_PBT, M1P_MJ, M1P_MJSW, M1P_FC,
       M1P_F2D, M1P_F3D, M1P_F4D, M1P_F2S, M1P_F3S, M1P_F4S, M1P_CGSO, M1P_CGDO, M1P_CGBO, M1P_SEED, M1P_K };
// QON: the raw CJO is positive (the depletion charge is on); RSON: the raw RS
// is positive (ai is its own row, behind RST = RS(T)).
enum { DIOP_IS = 0, DIOP_VTE, DIOP_GMIN, DIOP_VCRIT, DIOP_CJO, DIOP_VJ, DIOP_MT, DIOP_FC, DIOP_TT, DIOP_QON, DIOP_RSON,
       DIOP_RST, DIOP_K };

// Local unknowns, ports, user and prepared parameter counts per kind.
// The diode's unknowns are (a, c, ai): ai, the junction's anode behind RS,
// is the row of a itself when RS is 0 (diode.va:10-11, the collapse).
constexpr int KIND_NU[K_NKIND] = {3, 2, 2, 2, 3, 4, PSP103_NU, BSIM4VA_NU};
constexpr int KIND_NPORT[K_NKIND] = {2, 2, 2, 2, 2, 4, PSP103_NPORT, BSIM4VA_NPORT};
constexpr int KIND_UK[K_NKIND] = {VSRC_K, ISRC_K, RES_K, CAP_K, DIO_K, M1_K, PSP103_NPARAM, BSIM4VA_NPARAM};
constexpr int KIND_PK[K_NKIND] = {SRCP_K, SRCP_K, 2, 1, DIOP_K, M1P_K, PSP103_NPREP, BSIM4VA_NPREP};
// room for one instance's user or prepared parameters on the host (the
// hand-written kinds; a compiled kind sizes its own)
constexpr int PARAM_MAX = 128;
static_assert(VSRC_K <= PARAM_MAX && SRCP_K <= PARAM_MAX && DIO_K <= PARAM_MAX && DIOP_K <= PARAM_MAX &&
                  M1_K <= PARAM_MAX && M1P_K <= PARAM_MAX,
              "every kind's parameters fit PARAM_MAX");
// Residual rows a kind writes in a static evaluation.
constexpr unsigned KIND_RMASK[K_NKIND] = {0x7u, 0x3u, 0x3u, 0x0u, 0x7u, 0xdu, (1u << PSP103_NU) - 1u,
                                          (1u << BSIM4VA_NU) - 1u};
// Unknowns a kind's limiter writes (its limited image).
constexpr unsigned KIND_LIMW[K_NKIND] = {0u, 0u, 0u, 0u, 0x4u, 0xbu, PSP103_LIMW, BSIM4VA_LIMW};
// The local row whose charge is the kind's LTE site (one ddt per capacitor
// and per diode), or -1; a compiled kind has one site per ddt() instead.
constexpr int KIND_TAPE_ROW[K_NKIND] = {-1, -1, -1, 0, 2, -1, -1, -1};
// LTE charge sites per instance.
constexpr int KIND_NSITE[K_NKIND] = {0, 0, 0, 1, 1, 0, PSP103_NSITE, BSIM4VA_NSITE};

// Structural G-plane and C-plane entries per local row (bit masks over
// local columns). The pattern and slot tapes use their union.
ESP_HD unsigned kind_gmask(int kind, int ru)
{
    switch (kind) {
    case K_VSRC: return ru == 2 ? 0x3u : 0x4u;
    case K_RES: return 0x3u;
    case K_DIODE: return ru == 0 ? 0x5u : ru == 1 ? 0x6u : 0x7u;
    case K_MOS1: return ru == 1 ? 0x0u : 0xfu;
    case K_PSP103: return psp103_gmask[ru];
    case K_BSIM4: return bsim4va_gmask[ru];
    default: return 0u;
    }
}
ESP_HD unsigned kind_cmask(int kind, int ru)
{
    switch (kind) {
    case K_CAP: return 0x3u;
    case K_DIODE: return ru == 0 ? 0x0u : 0x6u;
    case K_MOS1: return 0xfu;
    case K_PSP103: return psp103_cmask[ru];
    case K_BSIM4: return bsim4va_cmask[ru];
    default: (void)ru; return 0u;
    }
}
ESP_HD unsigned kind_pmask(int kind, int ru) { return kind_gmask(kind, ru) | kind_cmask(kind, ru); }

// $temperature of every run: 27 degC.
constexpr double TEMPERATURE_K = 300.15;
constexpr double KB = 1.38064852e-23, QE = 1.6021766208e-19;

// Compiled kinds (va_models.h): the setup, its collapse flags and the
// collapse pairs. setup returns 0, or 1 + the index of a parameter outside
// its declared range.
inline int va_setup(int kind, const double *u, double *p, double gmin = VA_SIMPARAM_GMIN)
{
    switch (kind) {
    case K_PSP103: return psp103_setup(u, TEMPERATURE_K, gmin, p);
    case K_BSIM4: return bsim4va_setup(u, TEMPERATURE_K, gmin, p);
    default: return -1;
    }
}
inline int va_ncoll(int kind) { return kind == K_PSP103 ? PSP103_NCOLL : kind == K_BSIM4 ? BSIM4VA_NCOLL : 0; }
inline int va_coll_base(int kind)
{
    return kind == K_PSP103 ? PSP103_COLL_BASE : kind == K_BSIM4 ? BSIM4VA_COLL_BASE : 0;
}
inline const int *va_coll_pairs(int kind)
{
    return kind == K_PSP103 ? psp103_collapse : kind == K_BSIM4 ? bsim4va_collapse : nullptr;
}
inline const char *const *va_param_names(int kind)
{
    return kind == K_PSP103 ? psp103_param_names : kind == K_BSIM4 ? bsim4va_param_names : nullptr;
}
// aliasparam names of a compiled kind and the parameter each stands for
inline int va_nalias(int kind) { return kind == K_PSP103 ? PSP103_NALIAS : kind == K_BSIM4 ? BSIM4VA_NALIAS : 0; }
inline const char *const *va_alias_names(int kind)
{
    return kind == K_PSP103 ? psp103_alias_names : kind == K_BSIM4 ? bsim4va_alias_names : nullptr;
}
inline const int *va_alias_target(int kind)
{
    return kind == K_PSP103 ? psp103_alias_target : kind == K_BSIM4 ? bsim4va_alias_target : nullptr;
}

// The unknowns of one compiled instance: ports as given, internal unknowns
// collapsed (union-find over the pairs its prepared constants switch on,
// a set with a port takes the port's row, else its first unknown's) or
// fresh rows in unknown order. rows[u] < 0 asks for a fresh row; the
// caller allocates it; rows[u] = -2 - r takes unknown r's row (r < u).
// Returns 0, the setup's range error, or -2 for a collapse of two ports.
inline int va_collapse_rows(int kind, const double *u, const uint32_t *ports, int64_t *rows)
{
    const int nu = KIND_NU[kind], np = KIND_NPORT[kind];
    double *p = new double[KIND_PK[kind]];
    const int rc = va_setup(kind, u, p);
    int parent[32];
    for (int k = 0; k < nu; k++) parent[k] = k;
    auto find = [&](int a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
    };
    const int *pr = va_coll_pairs(kind);
    for (int k = 0; rc == 0 && k < va_ncoll(kind); k++) {
        if (p[va_coll_base(kind) + k] == 0.0) continue;
        int a = find(pr[2 * k]), b = find(pr[2 * k + 1]);
        if (a == b) continue;
        if (a < np && b < np) {
            delete[] p;
            return -2;  // a collapse between two ports
        }
        // the smaller unknown index is the root: ports come first
        if (a < b) parent[b] = a;
        else parent[a] = b;
    }
    delete[] p;
    for (int k = 0; k < nu; k++) {
        const int r = find(k);
        rows[k] = r < np ? (int64_t)ports[r] : (r == k ? -1 : -2 - r);
    }
    return rc;
}

struct DevOut {
    double res[4];
    double jac[4][4];
    double q[4];        // charges (transient only)
    double cjac[4][4];  // dQ/du (transient only)
};

ESP_HD double given(double v, double dflt) { return v == v ? v : dflt; }

// mos1.va:151-284: bias-independent defaulting, temperature, geometry.
inline void prep_mos1(const double *u, double *p)
{
    const double type = given(u[M1_TYPE], 1.0), tnom = given(u[M1_TNOM], 27.0);
    const double kp = given(u[M1_KP], 2.0e-5), phi = given(u[M1_PHI], 0.6), gamma = given(u[M1_GAMMA], 0.0);
    const double vto = given(u[M1_VTO], 0.0), tox = given(u[M1_TOX], 1.0e-7), u0 = given(u[M1_U0], 600.0);
    const double nsub = given(u[M1_NSUB], 0.0), nss = given(u[M1_NSS], 0.0), tpg = given(u[M1_TPG], 1.0);
    const double lambda = given(u[M1_LAMBDA], 0.0), ld = given(u[M1_LD], 0.0);
    const double is = given(u[M1_IS], 1.0e-14), js = given(u[M1_JS], 0.0);
    const double cbd = given(u[M1_CBD], 0.0), cbs = given(u[M1_CBS], 0.0), cj = given(u[M1_CJ], 0.0);
    const double cjsw = given(u[M1_CJSW], 0.0), pb = given(u[M1_PB], 0.8), mj = given(u[M1_MJ], 0.5);
    const double mjsw = given(u[M1_MJSW], 0.5), fc = given(u[M1_FC], 0.5);
    const double cgso = given(u[M1_CGSO], 0.0), cgdo = given(u[M1_CGDO], 0.0), cgbo = given(u[M1_CGBO], 0.0);
    const double gmin = given(u[M1_GMIN], 1.0e-12), dtemp = given(u[M1_DTEMP], 0.0);
    const double l = given(u[M1_L], 1.0e-4), w = given(u[M1_W], 1.0e-4);
    const double ad = given(u[M1_AD], 0.0), as = given(u[M1_AS], 0.0), pd = given(u[M1_PD], 0.0), ps = given(u[M1_PS], 0.0);
    const double m = given(u[M1_M], 1.0);
    const bool tox_given = u[M1_TOX] == u[M1_TOX], nsub_given = u[M1_NSUB] == u[M1_NSUB];
    const bool phi_given = u[M1_PHI] == u[M1_PHI], gamma_given = u[M1_GAMMA] == u[M1_GAMMA];
    const bool vto_given = u[M1_VTO] == u[M1_VTO], kp_given = u[M1_KP] == u[M1_KP];
    const bool temp_given = u[M1_TEMP] == u[M1_TEMP];

    double tnomk = tnom + 273.15;
    double vtnom = tnomk * (KB / QE);
    double egfet1 = 1.16 - (7.02e-4 * tnomk * tnomk) / (tnomk + 1108.0);
    double coxp = (tox_given && tox > 0.0) ? 3.9 * 8.854214871e-12 / tox : 0.0;
    double phix = phi, gammax = gamma, vtox = vto;
    if (nsub_given && nsub * 1.0e6 > 1.45e16) {
        if (!phi_given) {
            phix = 2.0 * vtnom * e_log(nsub * 1.0e6 / 1.45e16);
            if (phix < 0.1) phix = 0.1;
        }
        if (coxp > 0.0) {
            if (!gamma_given) gammax = __builtin_sqrt(2.0 * 11.7 * 8.854214871e-12 * 1.6021766208e-19 * nsub * 1.0e6) / coxp;
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
                vtox = vfb + type * (gammax * __builtin_sqrt(phix) + phix);
            }
        }
    }
    double kpx = (kp_given || coxp <= 0.0) ? kp : u0 * coxp * 1.0e-4;
    double tk = temp_given ? u[M1_TEMP] + 273.15 : TEMPERATURE_K + dtemp;
    double vt = tk * (KB / QE);
    double fact1 = tnomk / 300.15, fact2 = tk / 300.15;
    double egfet = 1.16 - (7.02e-4 * tk * tk) / (tk + 1108.0);
    double arg1 = -egfet1 / (2.0 * KB * tnomk) + 1.1150877 / (2.0 * KB * 300.15);
    double argt = -egfet / (2.0 * KB * tk) + 1.1150877 / (2.0 * KB * 300.15);
    double pbfact1 = -2.0 * vtnom * (1.5 * e_log(fact1) + QE * arg1);
    double pbfact = -2.0 * vt * (1.5 * e_log(fact2) + QE * argt);
    double ratio4 = (tk / tnomk) * __builtin_sqrt(tk / tnomk);
    double kpt = kpx / ratio4;
    double phio = (phix - pbfact1) / fact1;
    double phit = fact2 * phio + pbfact;
    double vbit = vtox - type * (gammax * __builtin_sqrt(phix)) + 0.5 * (egfet1 - egfet) + type * 0.5 * (phit - phix);
    double ist = is * e_exp(-egfet / vt + egfet1 / vtnom);
    double jst = js * e_exp(-egfet / vt + egfet1 / vtnom);
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
    double sargb = e_pow(argf, -mj), sargs = e_pow(argf, -mjsw);
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
    p[M1P_VCRITS] = vt * e_log(vt / (1.4142135623730951 * isats));
    p[M1P_VCRITD] = vt * e_log(vt / (1.4142135623730951 * isatd));
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
    p[M1P_SEED] = type * vbit + gammax * __builtin_sqrt(phit);
}

// diode.va:66-172, 259-262 (tlev = tlevc = 0, level 1, no sidewall).
inline void prep_diode(const double *u, double *p)
{
    const double is = given(u[DIO_IS], 1e-14), n = given(u[DIO_N], 1.0), cjo = given(u[DIO_CJO], 0.0);
    const double vj = given(u[DIO_VJ], 1.0), m = given(u[DIO_M], 0.5), fc = given(u[DIO_FC], 0.5);
    const double tt = given(u[DIO_TT], 0.0), gmin = given(u[DIO_GMIN], 1e-12), area = given(u[DIO_AREA], 1.0);
    const double tnom = given(u[DIO_TNOM], 27.0), dtemp = given(u[DIO_DTEMP], 0.0), xti = given(u[DIO_XTI], 3.0);
    const bool temp_given = u[DIO_TEMP] == u[DIO_TEMP];
    double tk = temp_given ? u[DIO_TEMP] + 273.15 : TEMPERATURE_K + dtemp;
    double tnomk = tnom + 273.15;
    double ratio = tk / tnomk, dtk = tk - tnomk;
    double vt = KB / QE * tk, vtnom = KB / QE * tnomk;
    double vte = n * vt;
    double ae = area;
    double egp = given(u[DIO_EG], 1.11);
    double egfet = 1.16 - (7.02e-4 * tk * tk) / (tk + 1108.0);
    double egfet1 = 1.16 - (7.02e-4 * tnomk * tnomk) / (tnomk + 1108.0);
    const double tm1 = given(u[DIO_TM1], 0.0), tm2 = given(u[DIO_TM2], 0.0);
    const double rs = given(u[DIO_RS], 0.0), trs = given(u[DIO_TRS], 0.0), trs2 = given(u[DIO_TRS2], 0.0);
    double mt = m * (1.0 + tm1 * dtk + tm2 * dtk * dtk);
    double fact1 = tnomk / 300.15, fact2 = tk / 300.15;
    double arg1 = -egfet1 / (2.0 * KB * tnomk) + 1.1150877 / (2.0 * KB * 300.15);
    double argt = -egfet / (2.0 * KB * tk) + 1.1150877 / (2.0 * KB * 300.15);
    double pbfact1 = -2.0 * vtnom * (1.5 * e_log(fact1) + QE * arg1);
    double pbfact = -2.0 * vt * (1.5 * e_log(fact2) + QE * argt);
    double pbo = (vj - pbfact1) / fact1;
    double gmaold = (vj - pbo) / pbo;
    double vj_t = pbfact + fact2 * pbo;
    double gmanew = (vj_t - pbo) / pbo;
    double cjo_t = cjo * (1.0 + mt * (400e-6 * (tk - 300.15) - gmanew)) / (1.0 + mt * (400e-6 * (tnomk - 300.15) - gmaold));
    cjo_t = cjo_t * ae;
    double is_t = is * ae * e_exp((ratio - 1.0) * egp / vte + (xti / n) * e_log(ratio));
    double issw_t = 0.0 * 0.0 * e_exp((ratio - 1.0) * egp / vt + (xti / 1.0) * e_log(ratio));
    is_t = is_t + issw_t;
    p[DIOP_IS] = is_t;
    p[DIOP_VTE] = vte;
    p[DIOP_GMIN] = gmin;
    p[DIOP_VCRIT] = vte * e_log(vte / (1.4142135623730951 * (is_t + 0.0)));
    p[DIOP_CJO] = cjo_t;
    p[DIOP_VJ] = vj_t;
    p[DIOP_MT] = mt;
    p[DIOP_FC] = fc;
    p[DIOP_TT] = tt;
    p[DIOP_QON] = cjo > 0.0 ? 1.0 : 0.0;
    p[DIOP_RSON] = rs > 0.0 ? 1.0 : 0.0;
    double rs_t = rs * (1.0 + trs * dtk + trs2 * dtk * dtk) / ae;
    if (rs_t < 1.0e-9) rs_t = 1.0e-9;
    p[DIOP_RST] = rs_t;
    (void)dtemp;
}

inline void prep_kind(int kind, const double *u, double *p)
{
    switch (kind) {
    case K_MOS1: prep_mos1(u, p); break;
    case K_DIODE: prep_diode(u, p); break;
    case K_VSRC:
    case K_ISRC: {
        // 0 everywhere but the -1 sentinels of TR, TF, PW, PER and PWL's r=,
        // and EXP's 1 ns time constants (models/vsource.va:58-96)
        for (int k = 0; k < SRCP_K; k++) {
            const bool sentinel = (k >= SRCP_TR && k <= SRCP_PER) || k == SRCP_PWL_R || k == SRCP_FFM;
            const double dflt = sentinel ? -1.0 : (k == SRCP_ETAU1 || k == SRCP_ETAU2) ? 1.0e-9
                                                : k == SRCP_FMDI ? 90.0 : 0.0;
            p[k] = u[k] == u[k] ? u[k] : dflt;
        }
        // an integer point count within the table
        const double len = p[SRCP_PWL_LEN];
        p[SRCP_PWL_LEN] = len >= 1.0 ? (len < (double)PWL_CAP ? __builtin_floor(len) : (double)PWL_CAP) : 0.0;
        // .op biases at dc, which the builder fills from the waveform's t = 0
        // value when the card gives none (models/vsource.va:216-221)
        if (u[VSRC_DC] != u[VSRC_DC]) {
            if (p[SRCP_WAVE] == 1.0) p[SRCP_DC] = p[SRCP_V1];
            else if (p[SRCP_WAVE] == 2.0) p[SRCP_DC] = p[SRCP_SVO] + p[SRCP_SVA] * e_sin(2.0 * VA_PI * p[SRCP_SPHASE] / 360.0);
            else if (p[SRCP_WAVE] == 3.0) p[SRCP_DC] = p[SRCP_EV1];
            else if (p[SRCP_WAVE] == 4.0) p[SRCP_DC] = p[SRCP_PWL_V];
        }
        break;
    }
    case K_RES:
        // resistor.va: rt = r at the nominal temperature (no tempcos), and
        // I(p, n) <+ V(p, n) / rt, whose slope is the reciprocal
        p[0] = u[RES_R];
        p[1] = 1.0 / u[RES_R];
        break;
    case K_CAP: p[0] = u[CAP_C]; break;
    default: break;
    }
}

// ------------------------------------------------------------- limiting
// fetlim, limvds, pnjlim and snap_limited: limiting.h (included through
// va_models.h, whose compiled kinds limit with them too).

// -------------------------------------------------------------- sources

ESP_HD double src_max(double a, double b) { return a > b ? a : b; }

// The PULSE widths of models/vsource.va:44-50 (P_TR, P_TF, P_PW, P_PER).
template <class P>
ESP_HD void pulse_widths(const P &p, double &tr, double &tf, double &pw, double &per)
{
    tr = src_max(p[SRCP_TR], 1.0e-12);
    tf = src_max(p[SRCP_TF], 1.0e-12);
    pw = src_max(p[SRCP_PW], 0.0) + src_max(-p[SRCP_PW], 0.0) * 1.0e30;
    per = src_max(p[SRCP_PER], tr + pw + tf + 1.0e-12);
}

// PWL (models/vsource.va:260-299): the table with r= repeat and td=
// delay, bisection for the segment, 1 fs the steepest ramp; an empty table is
// the DC value.
template <class P>
ESP_HD double pwl_value(const P &p, double t)
{
    const int len = (int)p[SRCP_PWL_LEN];
    if (len <= 0) return p[SRCP_DC];
    double te = t - p[SRCP_PWL_TD];
    const double t_first = p[SRCP_PWL_T], v_first = p[SRCP_PWL_V];
    const double t_last = p[SRCP_PWL_T + len - 1], v_last = p[SRCP_PWL_V + len - 1];
    const double r = p[SRCP_PWL_R];
    if (r >= 0.0 && len > 1 && te > t_last && t_last > r) {
        const double span = t_last - r;
        if (span > 0.0) {
            const double excess = te - r;
            te = r + (excess - __builtin_floor(excess / span) * span);
        }
    }
    if (te < 0.0 || te <= t_first) return v_first;
    if (te >= t_last) return v_last;
    int klo = 0, khi = len - 1;
    while (khi - klo > 1) {
        const int k = (klo + khi) / 2;
        if (p[SRCP_PWL_T + k] > te) khi = k;
        else klo = k;
    }
    const double tp = p[SRCP_PWL_T + klo], vp = p[SRCP_PWL_V + klo];
    const double tk = p[SRCP_PWL_T + khi], vk = p[SRCP_PWL_V + khi];
    return vp + (vk - vp) * (te - tp) / src_max(tk - tp, 1.0e-15);
}

// models/vsource.va:226-299: the DC value, or PULSE (PHASE 0), SIN or PWL
// at t.
template <class P>
ESP_HD double src_value(const P &p, double t)
{
    if (p[SRCP_WAVE] == 2.0) {
        // SIN (models/vsource.va:233-244): before TD the waveform's own t = 0
        // value; after it the damped sine
        const double t_sin = t - p[SRCP_STD];
        if (t_sin < 0.0) return p[SRCP_SVO] + p[SRCP_SVA] * e_sin(p[SRCP_SPHASE] * VA_PI / 180.0);
        return p[SRCP_SVO] + p[SRCP_SVA] * e_sin(VA_TWO_PI * p[SRCP_SFREQ] * t_sin + p[SRCP_SPHASE] * VA_PI / 180.0) *
                                 e_exp(va_min(-p[SRCP_STHETA] * t_sin, 80.0));
    }
    if (p[SRCP_WAVE] == 4.0) return pwl_value(p, t);
    if (p[SRCP_WAVE] == 5.0) {
        // SFFM (models/vsource.va:304-317): zero until TD, then the carrier
        // phase-modulated by the clamped index
        double mdi = p[SRCP_FMDI];
        if (mdi > p[SRCP_FFC] / p[SRCP_FFM]) mdi = p[SRCP_FFC] / p[SRCP_FFM];
        else if (mdi < 0.0) mdi = 0.0;
        const double t_fm = t - p[SRCP_FTD];
        if (t_fm <= 0.0) return 0.0;
        return p[SRCP_FVO] + p[SRCP_FVA] * e_sin((VA_TWO_PI * p[SRCP_FFC] * t_fm + p[SRCP_FPHC] * VA_PI / 180.0) +
                                                 mdi * e_sin(VA_TWO_PI * p[SRCP_FFM] * t_fm + p[SRCP_FPHM] * VA_PI / 180.0));
    }
    if (p[SRCP_WAVE] == 6.0) {
        // AM (vsource.va and isource.va alike): 0 before TD
        if (!(t >= p[SRCP_ATD])) return 0.0;
        const double env = p[SRCP_AVO] + e_sin(VA_TWO_PI * p[SRCP_AMF] * t + p[SRCP_APHS] * VA_PI / 180.0);
        const double car = e_sin(VA_TWO_PI * p[SRCP_AFC] * t + p[SRCP_APHC] * VA_PI / 180.0);
        return p[SRCP_AVA] * env * car;
    }
    if (p[SRCP_WAVE] == 3.0) {
        // EXP (models/vsource.va:252-258): two one-sided exponentials on V1
        const double t1 = t - p[SRCP_ETD1], t2 = t - p[SRCP_ETD2];
        const double v1 = p[SRCP_EV1], v2 = p[SRCP_EV2];
        double out = v1;
        if (t1 > 0.0) out = out + (v2 - v1) * (1.0 - e_exp(va_min(-t1 / src_max(p[SRCP_ETAU1], 1.0e-15), 80.0)));
        if (t2 > 0.0) out = out + (v1 - v2) * (1.0 - e_exp(va_min(-t2 / src_max(p[SRCP_ETAU2], 1.0e-15), 80.0)));
        return out;
    }
    if (p[SRCP_WAVE] != 1.0) return p[SRCP_DC];
    double tr, tf, pw, per;
    pulse_widths(p, tr, tf, pw, per);
    const double v1 = p[SRCP_V1], v2 = p[SRCP_V2];
    const double t_rel = t - p[SRCP_TD];
    double t_mod = t_rel + (0.0 / 360.0) * per;
    t_mod = t_mod - __builtin_floor(t_mod / per) * per;
    t_mod = src_max(t_mod, 0.0);
    if (t_rel < 0.0) return v1;
    if (t_mod < tr) return v1 + (v2 - v1) * (t_mod / tr);
    if (t_mod < tr + pw) return v2;
    if (t_mod < tr + pw + tf) return v2 - (v2 - v1) * ((t_mod - (tr + pw)) / tf);
    return v1;
}

// ngspice's PULSE breakpoint after tq rounded from the accepted time t_acc
// (src/device/eval.zig:1636-1650 spicePulseBreak); inf when none.
template <class P>
ESP_HD double pulse_break(const P &p, double t_acc, double tq)
{
    if (p[SRCP_WAVE] != 1.0) return inf_();
    double tr, tf, pw, per;
    pulse_widths(p, tr, tf, pw, per);
    double time = t_acc - p[SRCP_TD];
    if (time >= per) time -= per * __builtin_floor(time / per);
    const double atime = time + (tq - t_acc);
    const double corner = atime < 0.0 ? 0.0
                          : atime < tr ? tr
                          : atime < tr + pw ? tr + pw
                          : atime < tr + pw + tf ? tr + pw + tf
                          : per;
    const double bp = t_acc + (corner - time);
    return bp > tq ? bp : inf_();
}

// The fire time after t of `@(timer(start, period))` (LRM 5.10.5): start +
// k period for the first k strictly after t.
ESP_HD double timer_next(double start, double period, double t)
{
    if (start > t) return start;
    const double k = __builtin_floor((t - start) / period) + 1.0;
    const double next = start + k * period;
    return next > t ? next : start + (k + 1.0) * period;
}

// vsource's own nextBreakpoint, the fallback of srcBreak when ngspice's
// rounding names no corner (src/device/eval.zig:1592-1598): the earliest of
// the four PULSE corner timers of models/vsource.va:137-140 after t, for
// every waveform (a non-PULSE source parks TD at 1e30).
template <class P>
ESP_HD double pulse_timer_break(const P &p, double t)
{
    double tr, tf, pw, per;
    pulse_widths(p, tr, tf, pw, per);
    const double td = p[SRCP_TD];
    const double c[4] = {timer_next(td, per, t), timer_next(td + tr, per, t), timer_next(td + tr + pw, per, t),
                         timer_next(td + tr + pw + tf, per, t)};
    double best = c[0];
    for (int k = 1; k < 4; k++) best = c[k] < best ? c[k] : best;
    return best;
}

// ngspice's PWL breakpoint of a V source (src/device/eval.zig:1651-1697
// spicePwlBreak): the first table time after the request, rounded from the
// accepted time; folded by r=; inf past the end of a table that does not
// repeat. Returns 0 for null (ask the timers), else 1 with *bp set.
template <class P>
ESP_HD int pwl_break(const P &p, double t_acc, double tq, double *bp)
{
    const int n = (int)p[SRCP_PWL_LEN];
    const double t_last = p[SRCP_PWL_T + n - 1], r = p[SRCP_PWL_R];
    const bool repeats = r >= 0.0 && n > 1 && t_last > r;
    const double period = t_last - r;
    double time = t_acc - p[SRCP_PWL_TD];
    if (time > t_last) {
        if (!repeats) {
            *bp = inf_();
            return 1;
        }
        time -= r;
        time -= period * __builtin_floor(time / period);
        time += r;
    }
    const double atime = time + (tq - t_acc);
    double corner = inf_();
    bool found = false;
    for (int k = 0; k < n && !found; k++)
        if (p[SRCP_PWL_T + k] > atime) {
            corner = p[SRCP_PWL_T + k];
            found = true;
        }
    if (!found) {
        if (!repeats) {
            *bp = inf_();
            return 1;
        }
        for (int k = 0; k < n && !found; k++)
            if (p[SRCP_PWL_T + k] + period > atime) {
                corner = p[SRCP_PWL_T + k] + period;
                found = true;
            }
        if (!found) return 0;
    }
    const double b = t_acc + (corner - time);
    if (!(b > tq)) return 0;
    *bp = b;
    return 1;
}

// The V source's next breakpoint after tq (srcBreak, src/device/eval.zig:
// 1592-1598): ngspice's PULSE corner, else its PWL corner, else vsource's own
// timers (the PULSE corners, parked at 1e30 for other waveforms, and one
// one-shot timer per PWL slot, models/vsource.va:137-204), else inf.
template <class P>
ESP_HD double src_break(const P &p, double t_acc, double tq)
{
    const double pb = pulse_break(p, t_acc, tq);
    if (pb != inf_()) return pb;
    double b;
    if (p[SRCP_WAVE] == 4.0 && p[SRCP_PWL_LEN] > 0 && pwl_break(p, t_acc, tq, &b)) return b;
    double best = pulse_timer_break(p, tq);
    for (int k = 0; k < PWL_CAP; k++) {
        const double at = p[SRCP_PWL_T + k] + p[SRCP_PWL_TD];
        if (at > tq && at < best) best = at;
    }
    return best;
}

// ---------------------------------------------------------------- eval

ESP_HD Dual jcur(Dual v, double isat, double vtx, double gm)
{
    if (v.v > -3.0 * vtx) {
        Dual w = ddivs(v, vtx);
        Dual u = dminc(w, 80.0);
        Dual e = dexp(u);
        Dual t = dmul(e, dadd(dc(1.0), dsub(w, u)));
        return dadd(dscale(isat, daddc(t, -1.0)), dscale(gm, v));
    }
    return daddc(dscale(gm, v), -isat);
}

ESP_HD Dual sbody(Dual vbx, double phit)
{
    if (vbx.v <= 0.0) return dsqrt(dcsub(phit, vbx));
    double sp = __builtin_sqrt(phit);
    return dmaxc(dcsub(sp, ddivs(vbx, 2.0 * sp)), 0.0);
}

template <class P>
ESP_HD void eval_mos1(const P &p, const double *lx, DevOut &o)
{
    const double type = p[M1P_TYPE];
    Dual d = dvar(lx[0], 0), g = dvar(lx[1], 1), s = dvar(lx[2], 2), b = dvar(lx[3], 3);
    Dual vgs = dscale(type, dsub(g, s));
    Dual vds = dscale(type, dsub(d, s));
    Dual vbs = dscale(type, dsub(b, s));
    Dual vbd = dscale(type, dsub(b, d));
    Dual vgd = dsub(vgs, vds);
    Dual ibs = jcur(vbs, p[M1P_ISATS], p[M1P_VT], p[M1P_GMIN]);
    Dual ibd = jcur(vbd, p[M1P_ISATD], p[M1P_VT], p[M1P_GMIN]);
    double mode;
    Dual vbx, vgsx;
    if (vds.v >= 0.0) {
        mode = 1.0;
        vbx = vbs;
        vgsx = vgs;
    } else {
        mode = -1.0;
        vbx = vbd;
        vgsx = vgd;
    }
    Dual vdsx = dscale(mode, vds);
    Dual sarg = sbody(vbx, p[M1P_PHIT]);
    if (vbx.v > 0.0 && sarg.v > 0.0) {
        double sarg0 = sarg.v, vbx0 = vbx.v;
        sarg = dcsub(sarg0, ddivs(daddc(vbx, -vbx0), 2.0 * sarg0));
    }
    Dual von = dadd(dc(p[M1P_VBIT] * type), dscale(p[M1P_GAMMA], sarg));
    Dual vgst = dsub(vgsx, von);
    Dual bp = dscale(p[M1P_BETA], dadd(dc(1.0), dscale(p[M1P_LAMBDA], vdsx)));
    Dual id;
    if (vgst.v <= 0.0) id = dc(0.0);
    else if (vgst.v <= vdsx.v) id = dmul(dmul(dscale(0.5, bp), vgst), vgst);
    else id = dmul(dmul(bp, vdsx), dsub(vgst, dscale(0.5, vdsx)));
    Dual cd = dscale(type * mode, id);
    Dual tibd = dscale(type, ibd), tibs = dscale(type, ibs);
    Dual rd = dsub(cd, tibd);
    Dual rs = dsub(dneg(cd), tibs);
    Dual rb = dadd(tibd, tibs);
    o.res[0] = rd.v;
    o.res[1] = 0.0;
    o.res[2] = rs.v;
    o.res[3] = rb.v;
    for (int c = 0; c < 4; c++) {
        o.jac[0][c] = rd.d[c];
        o.jac[1][c] = 0.0;
        o.jac[2][c] = rs.d[c];
        o.jac[3][c] = rb.d[c];
    }
}

template <class P>
ESP_HD int limit_mos1(const P &p, const double *cur, const double *old, double *lim)
{
    const double type = p[M1P_TYPE];
    double vgs_o = type * (old[1] - old[2]), vds_o = type * (old[0] - old[2]);
    double vbs_o = type * (old[3] - old[2]), vbd_o = type * (old[3] - old[0]);
    double vgd_o = vgs_o - vds_o;
    double vgs = type * (cur[1] - cur[2]), vds = type * (cur[0] - cur[2]);
    double vbs = type * (cur[3] - cur[2]), vbd = type * (cur[3] - cur[0]);
    double vgd = vgs - vds;
    double vbx_o = vds_o >= 0.0 ? vbs_o : vbd_o;
    double phit = p[M1P_PHIT];
    double sb;
    if (vbx_o <= 0.0) sb = __builtin_sqrt(phit - vbx_o);
    else sb = va_max(__builtin_sqrt(phit) - vbx_o / (2.0 * __builtin_sqrt(phit)), 0.0);
    double von = p[M1P_VBIT] * type + p[M1P_GAMMA] * sb;
    if (vds_o >= 0.0) {
        vgs = fetlim(vgs, vgs_o, von);
        vds = vgs - vgd;
        vds = limvds(vds, vds_o);
    } else {
        vgd = fetlim(vgd, vgd_o, von);
        vds = vgs - vgd;
        vds = -limvds(-vds, -vds_o);
        vgs = vgd + vds;
    }
    int check = 0;
    if (vds >= 0.0) {
        vbs = pnjlim(vbs, vbs_o, p[M1P_VT], p[M1P_VCRITS], &check);
    } else {
        vbd = pnjlim(vbd, vbd_o, p[M1P_VT], p[M1P_VCRITD], &check);
        vbs = vbd + vds;
    }
    double s = cur[2];
    lim[0] = s + type * vds;
    lim[1] = s + type * vgs;
    lim[2] = s;
    lim[3] = s + type * vbs;
    snap_limited(4, 0xfu, 0xbu, lim, cur);
    return check;
}

template <class P>
ESP_HD void seed_mos1(const P &p, const double *cur, double *lim)
{
    const double type = p[M1P_TYPE];
    double s = cur[2];
    lim[0] = s;
    lim[1] = s + type * p[M1P_SEED];
    lim[2] = s;
    lim[3] = s + type * -1.0;
}

// diode.va:282-372 (no ikf/ikr/isr/bv, level 1): the junction current and
// charge across (ai, c) at the limited vd, and the series branch (a, ai),
// V/RS(T), when RS is on. Unknowns (a, c, ai); rows a: +ir, c: -ij,
// ai: ij - ir; charge rows c: -qj, ai: +qj.
template <class P>
ESP_HD void eval_diode(const P &p, const double *lx, DevOut &o, bool tran)
{
    const Dual a = dvar(lx[0], 0), c = dvar(lx[1], 1), ai = dvar(lx[2], 2);
    const Dual vd = dsub(ai, c);
    const double vte = p[DIOP_VTE], is_t = p[DIOP_IS];
    Dual id;
    if (vd.v >= -3.0 * vte) {
        Dual u = ddivs(vd, vte);
        Dual ulim = dminc(u, 80.0);
        id = dscale(is_t, daddc(dmul(dexp(ulim), dadd(dc(1.0), dsub(u, ulim))), -1.0));
    } else {
        Dual t = dpowc(ddiv(dc(3.0 * vte), dscale(2.718281828459045, vd)), 3.0);
        id = dscale(-is_t, dadd(dc(1.0), t));
    }
    const Dual ij = dadd(id, dscale(p[DIOP_GMIN], vd));
    const Dual ir = p[DIOP_RSON] != 0.0 ? ddivs(dsub(a, ai), p[DIOP_RST]) : dc(0.0);
    const Dual rai = dsub(ij, ir);
    o.res[0] = ir.v;
    o.res[1] = -ij.v;
    o.res[2] = rai.v;
    for (int k = 0; k < 3; k++) {
        o.jac[0][k] = ir.d[k];
        o.jac[1][k] = -ij.d[k];
        o.jac[2][k] = rai.d[k];
    }
    if (!tran) return;
    // one junction charge, depletion plus diffusion (diode.va:328-363)
    Dual qj = dc(0.0);
    if (p[DIOP_QON] != 0.0) {
        const double cjo_t = p[DIOP_CJO], vj_t = p[DIOP_VJ], mt = p[DIOP_MT], fc = p[DIOP_FC];
        const double vfc = fc * vj_t;
        if (vd.v < vfc) {
            qj = ddivs(dscale(cjo_t * vj_t, dcsub(1.0, dpowc(dcsub(1.0, ddivs(vd, vj_t)), 1.0 - mt))), 1.0 - mt);
        } else {
            const double f1 = vj_t * (1.0 - e_pow(1.0 - fc, 1.0 - mt)) / (1.0 - mt);
            const double f2 = e_pow(1.0 - fc, 1.0 + mt);
            const double f3 = 1.0 - fc * (1.0 + mt);
            const Dual lin = dscale(f3, daddc(vd, -vfc));
            const Dual quad = dscale(mt / (2.0 * vj_t), daddc(dmul(vd, vd), -(vfc * vfc)));
            qj = dscale(cjo_t, dadd(dc(f1), dscale(1.0 / f2, dadd(lin, quad))));
        }
    }
    qj = dadd(qj, dscale(0.0, vd));  // + ccox*vd, no plate caps at level 1
    if (p[DIOP_TT] > 0.0) qj = dadd(qj, dscale(p[DIOP_TT], ij));
    o.q[1] = -qj.v;
    o.q[2] = qj.v;
    for (int k = 0; k < 3; k++) {
        o.cjac[1][k] = -qj.d[k];
        o.cjac[2][k] = qj.d[k];
    }
}

// pnjlim on V(ai, c), written back as ai's image (reads c and ai).
template <class P>
ESP_HD int limit_diode(const P &p, const double *cur, const double *old, double *lim)
{
    int check = 0;
    double vd = pnjlim(cur[2] - cur[1], old[2] - old[1], p[DIOP_VTE], p[DIOP_VCRIT], &check);
    lim[0] = cur[0];
    lim[1] = cur[1];
    lim[2] = cur[1] + vd;
    snap_limited(3, 0x6u, 0x4u, lim, cur);
    return check;
}

template <class P>
ESP_HD void seed_diode(const P &p, const double *cur, double *lim)
{
    lim[0] = cur[0];
    lim[1] = cur[1];
    lim[2] = cur[1] + p[DIOP_VCRIT];
}

// Evaluation of one instance of `kind` at local voltages lx. In a
// transient (`tran`) sources read their waveform at t and the charges and
// dQ/du of capacitors are filled in (capacitor.va: charge c*V(p,n)).
// Independent sources multiply their value by `srcfact`, the operating
// point's source-stepping factor $simparam("sourceScaleFactor")
// (vsource.va:333, isource.va:332); 1 leaves the value's bits.
template <class P>
ESP_HD void eval_kind(int kind, const P &p, const double *lx, DevOut &o, bool tran = false, double t = 0.0,
                      double srcfact = 1.0)
{
    for (int r = 0; r < 4; r++) {
        o.res[r] = 0.0;
        o.q[r] = 0.0;
        for (int c = 0; c < 4; c++) {
            o.jac[r][c] = 0.0;
            o.cjac[r][c] = 0.0;
        }
    }
    switch (kind) {
    case K_VSRC:
        o.res[0] = lx[2];
        o.res[1] = -lx[2];
        o.res[2] = lx[0] - lx[1] - (tran ? src_value(p, t) : p[SRCP_DC]) * srcfact;
        o.jac[0][2] = 1.0;
        o.jac[1][2] = -1.0;
        o.jac[2][0] = 1.0;
        o.jac[2][1] = -1.0;
        break;
    case K_ISRC: {
        // I(p, n) <+ out: dc at an operating point, the waveform in a
        // transient (models/isource.va:216-330)
        const double out = (tran ? src_value(p, t) : p[SRCP_DC]) * srcfact;
        o.res[0] = out;
        o.res[1] = -out;
        break;
    }
    case K_RES: {
        // V(p, n) / rt: a quotient by a constant, its lanes scaled by the
        // reciprocal (models/resistor.va, eval.zig's Dual division)
        double i = (lx[0] - lx[1]) / p[0];
        o.res[0] = i;
        o.res[1] = -i;
        o.jac[0][0] = p[1];
        o.jac[0][1] = -p[1];
        o.jac[1][0] = -p[1];
        o.jac[1][1] = p[1];
        break;
    }
    case K_CAP:
        if (tran) {
            const double qv = p[0] * (lx[0] - lx[1]);
            o.q[0] = qv;
            o.q[1] = -qv;
            o.cjac[0][0] = p[0];
            o.cjac[0][1] = -p[0];
            o.cjac[1][0] = -p[0];
            o.cjac[1][1] = p[0];
        }
        break;
    case K_DIODE: eval_diode(p, lx, o, tran); break;
    case K_MOS1: eval_mos1(p, lx, o); break;
    default: break;
    }
}

}  // namespace esp
