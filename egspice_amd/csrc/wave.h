// The workgroup-per-lane Newton of the lane engine.
//
// One 64-thread workgroup owns one lane and keeps all of its state in LDS:
// node voltages, the G plane and residual, the limiter images, the lane's
// prepared device constants and its whole sparse direct solver, after a
// copy of the template's own tables (pattern, order, instance list, stamp
// tape), so the serial solver's dependent loads never leave the CU. A Newton
// tick splits into phases separated by workgroup barriers:
//   eval     thread t evaluates instances t, t+64, ... (forward-mode AD and
//            the limiting correction) into per-instance staging cells;
//   reduce   thread t sums planes entries t, t+64, ... over their staging
//            cells in the serial evaluation order (the host-built stamp
//            tape), so every sum is bitwise the serial one
//            (src/analysis/gpu/lane_color.zig's ordered hub reduction plays
//            the same role in the reference);
//   solve    thread 0 runs the lane's direct factor and solve in LDS;
//   limit    threads over instances; then the gates on thread 0.
// Every phase is an ESP_HD function of (tid, nthreads), so the host test
// harness replays the same phases thread by thread.
#pragma once

#include "engine.h"

namespace esp {

// One entry of the serial evaluation order: kind, instance, and where its
// stamps, prepared constants and limiter image live in the lane's block.
struct InstEntry {
    uint8_t kind;
    uint8_t pad[3];
    uint32_t id;
    uint32_t stage_res;  // first residual cell (NU cells)
    uint32_t stage_jac;  // first Jacobian cell (one per G entry)
    uint32_t prm;        // first prepared constant in the lane's block
    uint32_t lim;        // first limiter image word
    uint32_t stage_q;    // first charge cell (NU cells), transient only
    uint32_t stage_c;    // first dQ/du cell (one per C entry), transient only
    uint32_t tape;       // LTE charge site, or ESP_NONE
};

// Template-wide data every workgroup reads (device or host pointers).
struct WaveTemplate {
    Pattern P;
    const uint32_t *diag;
    const uint8_t *cur_row;
    uint32_t n_inst;
    const InstEntry *inst;
    const uint32_t *nodes_flat;  // NU words per instance, serial order
    const uint32_t *slots_flat;  // kslots words per instance, serial order
    const uint32_t *node_off;    // per instance: first word in nodes_flat
    const uint32_t *slot_off;    // per instance: first word in slots_flat
    const uint32_t *rptr, *ridx;  // residual row r sums stage cells ridx[rptr[r]..rptr[r+1])
    const uint32_t *gptr, *gidx;  // G slot s sums stage cells gidx[gptr[s]..gptr[s+1])
    const uint32_t *qptr, *qidx;  // charge row r (transient)
    const uint32_t *cptr, *cidx;  // C slot s (transient)
    uint32_t n_tape;              // LTE charge sites
    uint32_t n_stage;             // staging cells per lane
    uint32_t n_prm;               // prepared constants per lane
    uint32_t n_prm_lds;           // the leading ones held in LDS
    uint32_t n_lim;               // limiter words per lane
    uint32_t cap;                 // L+U capacity per lane
};

static_assert(sizeof(InstEntry) == 36, "InstEntry packs into nine u32 words");

// The template's tables read from a copy of its blob (at `base`, u32 words).
ESP_HD WaveTemplate wave_template(const WaveBlob &B, const uint32_t *base)
{
    WaveTemplate W;
    W.P = Pattern{B.n, B.nnz, base + B.col_ptr, base + B.row_idx, base + B.q, B.tri ? base + B.tri_pos : nullptr,
                  B.diag_fill, B.cap};
    W.diag = base + B.diag;
    W.cur_row = reinterpret_cast<const uint8_t *>(base + B.cur_row);
    W.n_inst = B.n_inst;
    W.inst = reinterpret_cast<const InstEntry *>(base + B.inst);
    W.nodes_flat = base + B.nodes;
    W.slots_flat = base + B.slots;
    W.node_off = base + B.node_off;
    W.slot_off = base + B.slot_off;
    W.rptr = base + B.rptr;
    W.ridx = base + B.ridx;
    W.gptr = base + B.gptr;
    W.gidx = base + B.gidx;
    W.qptr = base + B.qptr;
    W.qidx = base + B.qidx;
    W.cptr = base + B.cptr;
    W.cidx = base + B.cidx;
    W.n_tape = B.n_tape;
    W.n_stage = B.n_stage;
    W.n_prm = B.n_prm;
    W.n_prm_lds = B.n_prm_lds;
    W.n_lim = B.n_lim;
    W.cap = B.cap;
    return W;
}

// LDS offsets (in 8-byte words) of one lane's block, computed on the host.
struct WaveLayout {
    uint32_t x, x_old, dx, g, rhs, lim, stage, prm, lu_f64, lu_u32, lu_u8, flags;
    uint32_t dense;  // wave_lu.h's dense replay arrays, or ESP_NONE when n > 64
    uint32_t tran;   // the transient's arrays (wave_tran.h), or ESP_NONE
    uint32_t words;
    // set by the caller before wave_words: the lean layout of templates
    // with compiled kinds, whose lanes are long-running and LDS-bound (the
    // workgroups a CU holds). The serial factor's L/U values and index
    // lists (lx, ux, li, ui; 3 cap words) and a transient's charge-site
    // history (the tape ring and its currents, 5 n_tape words) live in the
    // lane's HBM scratch (lane_scratch_words) rather than LDS, and the
    // kernels read the template's tables from global memory instead of an
    // LDS copy. The dense replay keeps its own copy of the factor and of
    // the U column lists in LDS, so only the serial factor, the dense
    // rebuild and the step's end touch the HBM arrays.
    uint32_t lu_global;
};

// 8-byte words of one lane's HBM scratch for a lu_global layout: the L/U
// arrays, then the tape ring (4 n_tape) and its currents (n_tape).
ESP_HD size_t lane_scratch_words(uint32_t cap, uint32_t n_tape) { return 3 * (size_t)cap + 5 * (size_t)n_tape; }

// 8-byte words of a transient lane's extra arrays (wave_tran.h TranLane):
// q (n), cc (nnz + 1), a (nnz + 1), the current tape (n_tape), the charge
// ring (4 n) and its dynamic current (n), the tape ring (4 n_tape) and its
// current (n_tape) unless they live in HBM (lu_global), x_prev and x_trial
// (n each), and 32 step scalars.
ESP_HD size_t tran_words(uint32_t n, uint32_t nnz, uint32_t n_tape, bool lu_global = false)
{
    const size_t tape_hist = lu_global ? 0 : 5 * (size_t)n_tape;
    return (size_t)n + 2 * ((size_t)nnz + 1) + n_tape + 5 * (size_t)n + tape_hist + 2 * (size_t)n + 32;
}

// 4-byte words of LDS the template's tables take in front of a lane block:
// none in the lean layout, which reads them from global memory.
ESP_HD size_t table_lds_words(uint32_t blob_words, const WaveLayout &L) { return L.lu_global ? 0 : blob_words; }

// Largest pattern the dense wave replay takes: one pivot row per thread of
// the lane's 64-wide wavefront.
constexpr uint32_t DENSE_MAX_N = 64;

// 8-byte words of the dense replay arrays: F (n*n f64), the two row masks
// (2n u64), amap (n*n u16) and porig (n u32); with the solver's lists in
// HBM, also an LDS copy of the U column lists (at most n(n-1)/2 u32).
ESP_HD size_t dense_words(uint32_t n, bool lu_global = false)
{
    const size_t ucols = lu_global ? ((size_t)n * (n - 1) / 2 + 1) / 2 : 0;
    return (size_t)n * n + 2 * (size_t)n + ((size_t)n * n + 3) / 4 + ((size_t)n + 1) / 2 + ucols;
}

// The lane's block carved into typed views (stride 1).
struct WaveLane {
    double *x, *x_old, *dx, *g, *rhs, *lim, *stage;
    const double *prm;
    const double *gprm;  // the lane's whole constants in HBM (compiled kinds read theirs there)
    LaneLU lu;
    uint32_t *flags;  // 0 singular, 1 limited, 2 done, 3 bits, 4 ok, 5 scaled bits lo, 6 hi
    double *q, *cc, *tape;  // transient: row charges, C plane, charge per LTE site
    // operating-point homotopy (the ladder's stepping rungs): where the
    // diagonal gmin lands, row i at column gstamp[i] (ESP_NONE: nowhere)
    // and slot gstamp[n + i], or every diagonal when null; the independent
    // sources' scale factor
    const uint32_t *gstamp;
    double srcfact;
};

ESP_HD size_t wave_words(const WaveTemplate &W, WaveLayout &L, bool transient = false, size_t tran_extra = 0)
{
    const size_t n = W.P.n, nnz = W.P.nnz, cap = W.cap;
    size_t at = 0;
    L.x = (uint32_t)at; at += n;
    L.x_old = (uint32_t)at; at += n;
    L.dx = (uint32_t)at; at += n;
    L.g = (uint32_t)at; at += nnz + 1;
    L.rhs = (uint32_t)at; at += n + 1;
    L.lim = (uint32_t)at; at += W.n_lim;
    L.stage = (uint32_t)at; at += W.n_stage;
    L.prm = (uint32_t)at; at += W.n_prm_lds;
    // doubles of the solver: lx, ux (cap, unless lu_global), udiag, w, y,
    // rscale (n), vcopy (nnz), tri (3n), probe (3n)
    const size_t lucap = L.lu_global ? 0 : cap;
    L.lu_f64 = (uint32_t)at; at += 2 * lucap + 4 * n + nnz + 6 * n;
    // u32 words of the solver: cnt 4, pinv n, lp n+1, up n+1, li, ui (cap,
    // unless lu_global), prow, void_slots (nnz), flag..prev 7n
    size_t u32w = 4 + n + 2 * (n + 1) + 2 * lucap + 2 * nnz + 7 * n;
    L.lu_u32 = (uint32_t)at; at += (u32w + 1) / 2;
    // bytes: state 3, void_col n, scaled_pivot n
    size_t u8w = 3 + 2 * n;
    L.lu_u8 = (uint32_t)at; at += (u8w + 7) / 8;
    L.flags = (uint32_t)at; at += 4;  // 8 u32 flags
    if (n <= DENSE_MAX_N) {
        L.dense = (uint32_t)at;
        at += dense_words((uint32_t)n, L.lu_global != 0);
    } else {
        L.dense = ESP_NONE;
    }
    L.tran = ESP_NONE;
    if (transient) {
        L.tran = (uint32_t)at;
        at += tran_words((uint32_t)n, (uint32_t)nnz, W.n_tape, L.lu_global != 0) + tran_extra;
    }
    L.words = (uint32_t)at;
    return at;
}

ESP_HD WaveLane wave_lane(const WaveTemplate &W, const WaveLayout &L, double *block)
{
    const uint32_t n = W.P.n, nnz = W.P.nnz, cap = W.cap;
    WaveLane s;
    s.x = block + L.x;
    s.x_old = block + L.x_old;
    s.dx = block + L.dx;
    s.g = block + L.g;
    s.rhs = block + L.rhs;
    s.lim = block + L.lim;
    s.stage = block + L.stage;
    s.prm = block + L.prm;
    s.gprm = s.prm;
    s.gstamp = nullptr;
    s.srcfact = 1.0;
    const uint32_t lucap = L.lu_global ? 0 : cap;
    double *f = block + L.lu_f64;
    LaneLU &u = s.lu;
    u.lx = LV<double>{f, 1}; f += lucap;
    u.ux = LV<double>{f, 1}; f += lucap;
    u.udiag = LV<double>{f, 1}; f += n;
    u.w = LV<double>{f, 1}; f += n;
    u.y = LV<double>{f, 1}; f += n;
    u.rscale = LV<double>{f, 1}; f += n;
    u.vcopy = LV<double>{f, 1}; f += nnz;
    u.bp = LV<double>{f, 1}; f += n;
    u.mul = LV<double>{f, 1}; f += n;
    u.cv = LV<double>{f, 1}; f += n;
    u.px = LV<double>{f, 1}; f += n;
    u.pr = LV<double>{f, 1}; f += n;
    u.prows = LV<double>{f, 1};
    uint32_t *w = reinterpret_cast<uint32_t *>(block + L.lu_u32);
    u.cnt = LV<uint32_t>{w, 1}; w += 4;
    u.pinv = LV<uint32_t>{w, 1}; w += n;
    u.lp = LV<uint32_t>{w, 1}; w += n + 1;
    u.up = LV<uint32_t>{w, 1}; w += n + 1;
    u.li = LV<uint32_t>{w, 1}; w += lucap;
    u.ui = LV<uint32_t>{w, 1}; w += lucap;
    u.prow = LV<uint32_t>{w, 1}; w += nnz;
    u.void_slots = LV<uint32_t>{w, 1}; w += nnz;
    u.flag = LV<uint32_t>{w, 1}; w += n;
    u.topo = LV<uint32_t>{w, 1}; w += n;
    u.stack = LV<uint32_t>{w, 1}; w += n;
    u.pstack = LV<uint32_t>{w, 1}; w += n;
    u.want = LV<uint32_t>{w, 1}; w += n;
    u.want_of = LV<uint32_t>{w, 1}; w += n;
    u.prev = LV<uint32_t>{w, 1};
    uint8_t *b = reinterpret_cast<uint8_t *>(block + L.lu_u8);
    u.state = LV<uint8_t>{b, 1}; b += 3;
    u.void_col = LV<uint8_t>{b, 1}; b += n;
    u.scaled_pivot = LV<uint8_t>{b, 1};
    s.flags = reinterpret_cast<uint32_t *>(block + L.flags);
    if (L.lu_global) u.lx = u.ux = LV<double>{nullptr, 1}, u.li = u.ui = LV<uint32_t>{nullptr, 1};
    s.q = s.cc = s.tape = nullptr;
    if (L.tran != ESP_NONE) {
        double *t = block + L.tran;
        s.q = t;
        s.cc = t + n;
        s.tape = t + n + 2 * ((size_t)nnz + 1);
    }
    return s;
}

// Points a lu_global lane's serial L/U arrays at its HBM scratch.
ESP_HD void wave_bind_lu_scratch(WaveLane &s, const WaveTemplate &W, double *scratch)
{
    s.lu.lx = LV<double>{scratch, 1};
    s.lu.ux = LV<double>{scratch + W.cap, 1};
    uint32_t *w = reinterpret_cast<uint32_t *>(scratch + 2 * (size_t)W.cap);
    s.lu.li = LV<uint32_t>{w, 1};
    s.lu.ui = LV<uint32_t>{w + W.cap, 1};
}

struct LP {
    const double *p;
    ESP_HD double operator[](int k) const { return p[k]; }
};

// The per-instance phases are instantiated per device kind so that the
// port count and the stamp masks are compile-time constants: every local
// array then lives in registers (a runtime port count would index them
// dynamically and push them to scratch memory).

// Phase: evaluate instance e into its staging cells.
template <int KIND, bool TRAN>
ESP_HD void wave_eval_k(const WaveTemplate &W, WaveLane &s, const InstEntry &ie, uint32_t e, bool lim_active, double t)
{
    constexpr int nu = KIND_NU[KIND];
    constexpr unsigned limw = KIND_LIMW[KIND];
    const uint32_t *nd = W.nodes_flat + W.node_off[e];
    LP p{s.prm + ie.prm};
    double lx[nu], xl[nu];
#pragma unroll
    for (int u = 0; u < nu; u++) xl[u] = lx[u] = s.x[nd[u]];
    const bool use_lim = limw != 0 && lim_active;
    if (use_lim) {
#pragma unroll
        for (int u = 0; u < nu; u++)
            if (limw >> u & 1) lx[u] = s.lim[ie.lim + u];
    }
    DevOut o;
    eval_kind(KIND, p, lx, o, TRAN, t, s.srcfact);
    bool charged = false;
#pragma unroll
    for (int ru = 0; ru < nu; ru++) charged = charged || kind_cmask(KIND, ru) != 0;
    if (use_lim) {
        // x minus its limited image, dotted with each row's gradient, on the
        // residual and (transient) on the charges (eval.zig:825-845, 893-897)
        double corr[nu];
        bool live = false;
#pragma unroll
        for (int u = 0; u < nu; u++) {
            corr[u] = xl[u] - lx[u];
            live = live || corr[u] != 0;
        }
        if (live) {
#pragma unroll
            for (int ru = 0; ru < nu; ru++) {
                double acc = 0;
#pragma unroll
                for (int u = 0; u < nu; u++) acc += o.jac[ru][u] * corr[u];
                o.res[ru] += acc;
            }
            if (TRAN && charged) {
#pragma unroll
                for (int ru = 0; ru < nu; ru++) {
                    if (kind_cmask(KIND, ru) == 0) continue;
                    double acc = 0;
#pragma unroll
                    for (int u = 0; u < nu; u++) acc += o.cjac[ru][u] * corr[u];
                    o.q[ru] += acc;
                }
            }
        }
    }
    uint32_t k = 0;
#pragma unroll
    for (int ru = 0; ru < nu; ru++) {
        s.stage[ie.stage_res + ru] = o.res[ru];
        const unsigned gm = kind_gmask(KIND, ru);
#pragma unroll
        for (int cu = 0; cu < nu; cu++)
            if (gm >> cu & 1) s.stage[ie.stage_jac + k++] = o.jac[ru][cu];
    }
    if (TRAN && charged) {
        uint32_t kc = 0;
#pragma unroll
        for (int ru = 0; ru < nu; ru++) {
            s.stage[ie.stage_q + ru] = o.q[ru];
            const unsigned cm = kind_cmask(KIND, ru);
#pragma unroll
            for (int cu = 0; cu < nu; cu++)
                if (cm >> cu & 1) s.stage[ie.stage_c + kc++] = o.cjac[ru][cu];
        }
        if (ie.tape != ESP_NONE) s.tape[ie.tape] = o.q[KIND_TAPE_ROW[KIND] < 0 ? 0 : KIND_TAPE_ROW[KIND]];
    }
}

// A compiled kind (va_models.h): the generated evaluation writes the
// instance's staging cells in the order the stamp tape reads them (rows,
// then the G mask's entries row-major; transient: charges, the C mask's
// entries, and one charge per ddt() site into the lane's charge tape).
template <int KIND, bool TRAN>
ESP_HD void va_eval_kind(const double *prm, const double *x, double *res, double *jac, double *q, double *cj,
                         double *site)
{
    if constexpr (KIND == K_PSP103) {
        if (TRAN) psp103_eval_tran(prm, x, res, jac, q, cj, site);
        else psp103_eval(prm, x, res, jac);
    } else if constexpr (KIND == K_BSIM4) {
        if (TRAN) bsim4va_eval_tran(prm, x, res, jac, q, cj, site);
        else bsim4va_eval(prm, x, res, jac);
    }
}

// A compiled kind's evaluation into its staging cells. With limiting active
// it runs at the limited image of the unknowns its limiter writes and, as
// for the hand-written kinds, adds x minus the image dotted with each row's
// gradient to the residual and (transient) the charge rows; the charge
// sites stay uncorrected (only the re-read at a solution, unlimited, feeds
// the step control).
template <int KIND, bool TRAN>
ESP_HD void wave_eval_va(const WaveTemplate &W, WaveLane &s, const InstEntry &ie, uint32_t e, bool lim_active)
{
    constexpr int nu = KIND_NU[KIND];
    constexpr unsigned limw = KIND_LIMW[KIND];
    const uint32_t *nd = W.nodes_flat + W.node_off[e];
    double x[nu];
#pragma unroll
    for (int u = 0; u < nu; u++) x[u] = s.x[nd[u]];
    const double *prm = s.gprm + ie.prm;
    double *res = s.stage + ie.stage_res, *jac = s.stage + ie.stage_jac;
    double *q = s.stage + ie.stage_q, *cj = s.stage + ie.stage_c;
    if constexpr (limw != 0) {
        if (lim_active) {
            double corr[nu];
            bool live = false;
#pragma unroll
            for (int u = 0; u < nu; u++) {
                const double xi = x[u];
                if (limw >> u & 1) x[u] = s.lim[ie.lim + u];
                corr[u] = xi - x[u];
                live = live || corr[u] != 0;
            }
            va_eval_kind<KIND, TRAN>(prm, x, res, jac, q, cj, s.tape + ie.tape);
            if (live) {
                uint32_t k = 0;
#pragma unroll
                for (int ru = 0; ru < nu; ru++) {
                    const unsigned gm = kind_gmask(KIND, ru);
                    double acc = 0;
                    for (int cu = 0; cu < nu; cu++)
                        if (gm >> cu & 1) acc += jac[k++] * corr[cu];
                    res[ru] += acc;
                }
                if (TRAN) {
                    uint32_t kc = 0;
#pragma unroll
                    for (int ru = 0; ru < nu; ru++) {
                        const unsigned cm = kind_cmask(KIND, ru);
                        double acc = 0;
                        for (int cu = 0; cu < nu; cu++)
                            if (cm >> cu & 1) acc += cj[kc++] * corr[cu];
                        q[ru] += acc;
                    }
                }
            }
            return;
        }
    }
    va_eval_kind<KIND, TRAN>(prm, x, res, jac, q, cj, s.tape + ie.tape);
}

// A compiled kind's charge rows and charge sites alone, values only (the
// charge re-read at a transient solution): the value path of the transient
// evaluation, so the same bits, without its derivative lanes or currents.
template <int KIND>
ESP_HD void wave_charge_va(const WaveTemplate &W, WaveLane &s, const InstEntry &ie, uint32_t e)
{
    constexpr int nu = KIND_NU[KIND];
    const uint32_t *nd = W.nodes_flat + W.node_off[e];
    double x[nu];
#pragma unroll
    for (int u = 0; u < nu; u++) x[u] = s.x[nd[u]];
    const double *prm = s.gprm + ie.prm;
    if constexpr (KIND == K_PSP103) psp103_eval_charge(prm, x, s.stage + ie.stage_q, s.tape + ie.tape);
    else if constexpr (KIND == K_BSIM4) bsim4va_eval_charge(prm, x, s.stage + ie.stage_q, s.tape + ie.tape);
}

// TRAN: sources read their waveform at t and charges are staged too. VA:
// the template holds compiled kinds; only then are their (long, register-
// hungry) evaluations instantiated, so the kernels of templates without
// them keep their own register budget and occupancy. QONLY (transient):
// only the charge rows and charge sites are read afterwards, so compiled
// kinds stage just those (the other kinds evaluate as usual).
template <bool TRAN, int VA, bool QONLY>
ESP_HD void wave_eval_inst_body(const WaveTemplate &W, WaveLane &s, uint32_t e, bool lim_active, double t)
{
    static_assert(TRAN || !QONLY, "charges are staged in a transient");
    const InstEntry ie = W.inst[e];
    if constexpr ((VA & VA_PSP) != 0) {
        if (ie.kind == K_PSP103) {
            if constexpr (QONLY) wave_charge_va<K_PSP103>(W, s, ie, e);
            else wave_eval_va<K_PSP103, TRAN>(W, s, ie, e, lim_active);
            return;
        }
    }
    if constexpr ((VA & VA_B4) != 0) {
        if (ie.kind == K_BSIM4) {
            if constexpr (QONLY) wave_charge_va<K_BSIM4>(W, s, ie, e);
            else wave_eval_va<K_BSIM4, TRAN>(W, s, ie, e, lim_active);
            return;
        }
    }
    switch (ie.kind) {
    case K_VSRC: wave_eval_k<K_VSRC, TRAN>(W, s, ie, e, lim_active, t); break;
    case K_ISRC: wave_eval_k<K_ISRC, TRAN>(W, s, ie, e, lim_active, t); break;
    case K_RES: wave_eval_k<K_RES, TRAN>(W, s, ie, e, lim_active, t); break;
    case K_CAP: wave_eval_k<K_CAP, TRAN>(W, s, ie, e, lim_active, t); break;
    case K_DIODE: wave_eval_k<K_DIODE, TRAN>(W, s, ie, e, lim_active, t); break;
    case K_MOS1: wave_eval_k<K_MOS1, TRAN>(W, s, ie, e, lim_active, t); break;
    default: break;
    }
}

// With compiled kinds the per-instance evaluation is one outlined call: the
// kernel around it (the solve, the step control) keeps its registers, and
// the call's register budget is the evaluations' alone.
template <bool TRAN, int VA, bool QONLY>
__host__ __device__ __attribute__((noinline)) void wave_eval_inst_call(const WaveTemplate &W, WaveLane &s, uint32_t e,
                                                                       bool lim_active, double t)
{
    wave_eval_inst_body<TRAN, VA, QONLY>(W, s, e, lim_active, t);
}

template <bool TRAN = false, int VA = 0, bool QONLY = false>
ESP_HD void wave_eval_inst(const WaveTemplate &W, WaveLane &s, uint32_t e, bool lim_active, double t = 0.0)
{
    if constexpr (VA != 0) wave_eval_inst_call<TRAN, VA, QONLY>(W, s, e, lim_active, t);
    else wave_eval_inst_body<TRAN, VA, QONLY>(W, s, e, lim_active, t);
}

// Phase (transient): charge row r (e < n) or C slot e - n over its staging
// cells in tape order, loaded eight at a time before the adds.
ESP_HD void wave_reduce_charge(const WaveTemplate &W, WaveLane &s, uint32_t e)
{
    const uint32_t n = W.P.n;
    const uint32_t *ptr = e < n ? W.qptr + e : W.cptr + (e - n);
    const uint32_t *idx = e < n ? W.qidx : W.cidx;
    const uint32_t c0 = ptr[0], c1 = ptr[1];
    double acc = 0.0;
    for (uint32_t c = c0; c < c1; c += 8) {
        uint32_t ix[8];
        double v[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; k++) ix[k] = idx[c + k < c1 ? c + k : c1 - 1];
#pragma unroll
        for (uint32_t k = 0; k < 8; k++) v[k] = s.stage[ix[k]];
#pragma unroll
        for (uint32_t k = 0; k < 8; k++)
            if (c + k < c1) acc += v[k];
    }
    if (e < n) s.q[e] = acc;
    else s.cc[e - n] = acc;
}

// Phase: residual row or G slot `e` of the planes (e < n rows, then slots):
// the sum over its staging cells in tape order. The cell indices, then the
// cells, are loaded eight at a time before the adds.
ESP_HD void wave_reduce_entry(const WaveTemplate &W, WaveLane &s, uint32_t e)
{
    const uint32_t n = W.P.n;
    const uint32_t *ptr = e < n ? W.rptr + e : W.gptr + (e - n);
    const uint32_t *idx = e < n ? W.ridx : W.gidx;
    const uint32_t c0 = ptr[0], c1 = ptr[1];
    double acc = 0.0;
    for (uint32_t c = c0; c < c1; c += 8) {
        uint32_t ix[8];
        double v[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; k++) ix[k] = idx[c + k < c1 ? c + k : c1 - 1];
#pragma unroll
        for (uint32_t k = 0; k < 8; k++) v[k] = s.stage[ix[k]];
#pragma unroll
        for (uint32_t k = 0; k < 8; k++)
            if (c + k < c1) acc += v[k];
    }
    if (e < n) s.rhs[e] = acc;
    else s.g[e - n] = acc;
}

// The reduce phase with the stamp tape held in registers: thread t's first
// entry (t < n + nnz) and up to RCACHE of its staging cells are read from
// the tape once per launch; each tick then issues those cell loads back to
// back and adds them in tape order. Entries past the first, or with more
// cells, take wave_reduce_entry.
constexpr uint32_t RCACHE = 16;

struct ReduceCache {
    uint32_t cnt;  // cells of the thread's first entry, or ESP_NONE when uncached
    uint32_t idx[RCACHE];
};

ESP_HD ReduceCache wave_reduce_cache(const WaveTemplate &W, uint32_t tid)
{
    ReduceCache c;
    c.cnt = ESP_NONE;
    const uint32_t n = W.P.n, e = tid;
    for (uint32_t k = 0; k < RCACHE; k++) c.idx[k] = 0;
    if (e >= n + W.P.nnz) return c;
    const uint32_t *ptr = e < n ? W.rptr + e : W.gptr + (e - n);
    const uint32_t *idx = e < n ? W.ridx : W.gidx;
    const uint32_t c0 = ptr[0], c1 = ptr[1];
    if (c1 - c0 > RCACHE) return c;
    c.cnt = c1 - c0;
    for (uint32_t k = 0; k < RCACHE; k++) c.idx[k] = k < c.cnt ? idx[c0 + k] : (c.cnt ? idx[c0] : 0);
    return c;
}

ESP_HD void wave_reduce(const WaveTemplate &W, WaveLane &s, const ReduceCache &c, uint32_t tid, uint32_t nt)
{
    const uint32_t n = W.P.n, total = n + W.P.nnz;
    uint32_t e = tid;
    if (c.cnt != ESP_NONE) {
        double v[RCACHE];
#pragma unroll
        for (uint32_t k = 0; k < RCACHE; k++) v[k] = s.stage[c.idx[k]];
        double acc = 0.0;
#pragma unroll
        for (uint32_t k = 0; k < RCACHE; k++)
            if (k < c.cnt) acc += v[k];
        if (e < n) s.rhs[e] = acc;
        else s.g[e - n] = acc;
        e += nt;
    }
    for (; e < total; e += nt) wave_reduce_entry(W, s, e);
}

// Phase: limit instance e (cur = x, old = image or x_old); true if clamped.
template <int KIND>
ESP_HD bool wave_limit_k(const WaveTemplate &W, WaveLane &s, const InstEntry &ie, uint32_t e, bool lim_active)
{
    constexpr int nu = KIND_NU[KIND];
    constexpr unsigned limw = KIND_LIMW[KIND];
    const uint32_t *nd = W.nodes_flat + W.node_off[e];
    double cur[nu], old[nu], lim[nu];
#pragma unroll
    for (int u = 0; u < nu; u++) {
        cur[u] = s.x[nd[u]];
        old[u] = ((limw >> u & 1) && lim_active) ? s.lim[ie.lim + u] : s.x_old[nd[u]];
    }
    int ch;
    if constexpr (KIND == K_BSIM4) {
        ch = bsim4va_limit(s.gprm + ie.prm, cur, old, lim);
    } else {
        LP p{s.prm + ie.prm};
        if constexpr (KIND == K_MOS1) ch = limit_mos1(p, cur, old, lim);
        else ch = limit_diode(p, cur, old, lim);
    }
#pragma unroll
    for (int u = 0; u < nu; u++)
        if (limw >> u & 1) s.lim[ie.lim + u] = lim[u];
    return ch != 0;
}

// VA: the compiled kinds present (their limiters are instantiated only
// then, as their evaluations are).
template <int VA = 0>
ESP_HD bool wave_limit_inst(const WaveTemplate &W, WaveLane &s, uint32_t e, bool lim_active)
{
    const InstEntry ie = W.inst[e];
    switch (ie.kind) {
    case K_DIODE: return wave_limit_k<K_DIODE>(W, s, ie, e, lim_active);
    case K_MOS1: return wave_limit_k<K_MOS1>(W, s, ie, e, lim_active);
    case K_BSIM4:
        if constexpr ((VA & VA_B4) != 0) return wave_limit_k<K_BSIM4>(W, s, ie, e, lim_active);
        return false;
    default: return false;
    }
}

template <int KIND>
ESP_HD void wave_seed_k(const WaveTemplate &W, WaveLane &s, const InstEntry &ie, uint32_t e)
{
    constexpr int nu = KIND_NU[KIND];
    constexpr unsigned limw = KIND_LIMW[KIND];
    const uint32_t *nd = W.nodes_flat + W.node_off[e];
    double cur[nu], lim[nu];
#pragma unroll
    for (int u = 0; u < nu; u++) cur[u] = s.x[nd[u]];
    if constexpr (KIND == K_BSIM4) {
        bsim4va_seed(s.gprm + ie.prm, cur, lim);
    } else {
        LP p{s.prm + ie.prm};
        if constexpr (KIND == K_MOS1) seed_mos1(p, cur, lim);
        else seed_diode(p, cur, lim);
    }
#pragma unroll
    for (int u = 0; u < nu; u++)
        if (limw >> u & 1) s.lim[ie.lim + u] = lim[u];
}

template <int VA = 0>
ESP_HD void wave_seed_inst(const WaveTemplate &W, WaveLane &s, uint32_t e)
{
    const InstEntry ie = W.inst[e];
    switch (ie.kind) {
    case K_DIODE: wave_seed_k<K_DIODE>(W, s, ie, e); break;
    case K_MOS1: wave_seed_k<K_MOS1>(W, s, ie, e); break;
    case K_BSIM4:
        if constexpr ((VA & VA_B4) != 0) wave_seed_k<K_BSIM4>(W, s, ie, e);
        break;
    default: break;
    }
}

// Fresh solver state (direct.Solver right after init).
ESP_HD void wave_fresh_solver(const WaveTemplate &W, WaveLane &s)
{
    LaneLU &u = s.lu;
    u.state[ST_FACTORED] = 0;
    u.state[ST_TRI] = W.P.tri_pos != nullptr ? 1 : 0;
    u.state[ST_LU_FACTORED] = 0;
    for (int i = 0; i < 4; i++) u.cnt[i] = 0;
    for (uint32_t i = 0; i < W.P.n; i++) u.w[i] = 0.0;
}

// Thread 0's finish of the planes: ground stamp, gmin, and at tick 0 the
// runaway bound (converger.zig loadExtras/plausibleLimit). The gmin lands
// on every diagonal, or at the lane's stamps (converger.zig:700-718
// loadGmin: vals[slot] += gmin, rhs[row] += gmin * x[col]).
ESP_HD void wave_assemble(const WaveTemplate &W, WaveLane &s, const Newton &o, uint32_t tick, double &limit)
{
    const uint32_t n = W.P.n;
    LV<double> g{s.g, 1}, rhs{s.rhs, 1}, x{s.x, 1};
    g[W.diag[0]] += 1.0;
    rhs[0] += x[0];
    if (o.gmin > 0) {
        if (s.gstamp != nullptr) {
            for (uint32_t i = 0; i < n; i++)
                if (s.gstamp[i] != ESP_NONE) g[s.gstamp[n + i]] += o.gmin;
            for (uint32_t i = 0; i < n; i++)
                if (s.gstamp[i] != ESP_NONE) rhs[i] += o.gmin * x[s.gstamp[i]];
        } else {
            for (uint32_t i = 0; i < n; i++) g[W.diag[i]] += o.gmin;
            for (uint32_t i = 0; i < n; i++) rhs[i] += o.gmin * x[i];
        }
    }
    if (tick == 0) limit = plausible_limit(n, x, rhs, g, W.diag);
}

// Thread 0's part of a tick after the planes: ground stamp, gmin, the
// runaway bound, factor and solve, update and score. Returns false when
// the lane is singular.
ESP_HD bool wave_solve(const WaveTemplate &W, WaveLane &s, const Newton &o, uint32_t tick, double &limit, double &scaled)
{
    const uint32_t n = W.P.n;
    LV<double> g{s.g, 1}, rhs{s.rhs, 1}, x{s.x, 1}, x_old{s.x_old, 1}, dx{s.dx, 1};
    wave_assemble(W, s, o, tick, limit);
    if (direct_factor(W.P, s.lu, g) != 0) return false;
    direct_solve_neg(W.P, s.lu, rhs, dx);
    scaled = update_norm(n, x, dx, x_old, W.cur_row, o.reltol, o.abstol, o.vntol);
    return true;
}

// The gates after the limit pass (finalizeStep's order), thread 0.
// `vals` is the matrix the tick factored (G, or G + ag0 C under the
// transient's hook): the residual gate scales rows by its diagonal.
ESP_HD bool wave_gates(const WaveTemplate &W, WaveLane &s, const Newton &o, uint32_t tick, bool limited, double limit,
                       double scaled, double *vals = nullptr)
{
    const uint32_t n = W.P.n;
    LV<double> g{vals ? vals : s.g, 1}, rhs{s.rhs, 1}, x{s.x, 1};
    bool ok = !limited && tick != 0 && !(scaled >= 1.0);
    if (ok) ok = plausible(n, x, limit);
    if (ok) ok = residual_ok(n, o, x, rhs, g, W.diag);
    return ok;
}

}  // namespace esp
