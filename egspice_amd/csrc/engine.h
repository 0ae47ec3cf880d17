// Internal declarations of the lane engine: the circuit template, the
// device-side views of a lane batch, and the per-lane static Newton body
// that the gfx950 kernel runs (one thread per lane).
#pragma once

#include <string>
#include <unordered_map>
#include <vector>

#include "devices.h"
#include "lane_solver.h"

namespace esp {

// -------------------------------------------------------------- ordering
uint32_t tarjan_blocks(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *emit, uint32_t *block_ptr);
uint64_t chol_count(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const uint32_t *perm, uint64_t cap);
int amd_order(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *q);
int column_order(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *q);

// ------------------------------------------------------------- template

// One netlist's frozen structure, shared by all lanes of a batch. Node rows
// are allocated on first touch (src/frontend/builder.zig:559-565 rowOf);
// a device's internal unknowns take fresh rows when it is added
// (builder.zig:238-266); kinds evaluate in first-use order, one batch per
// kind (src/device/Circuit.zig:112-201).
struct Template {
    uint32_t n = 1;  // row 0 is ground
    std::unordered_map<std::string, uint32_t> names;
    std::vector<uint8_t> cur_flag{0};
    std::vector<uint32_t> nodes[K_NKIND];
    std::vector<double> uprm[K_NKIND];
    uint32_t count[K_NKIND] = {0, 0, 0, 0, 0, 0};
    int order[K_NKIND] = {0, 0, 0, 0, 0, 0};
    int n_order = 0;
    bool frozen = false;
    // frozen structure
    uint32_t nnz = 0;
    std::vector<uint32_t> col_ptr, row_idx, diag, q, tri_pos, slots[K_NKIND];
    std::vector<uint8_t> cur_row;
    uint32_t kslots[K_NKIND] = {0, 0, 0, 0, 0, 0};
    uint64_t diag_fill = 0;
    bool tridiagonal = false;

    uint32_t node(const std::string &name);
    int add(int kind, const uint32_t *ports, const double *params, uint32_t *inst);
    int freeze();
    uint32_t find_slot(uint32_t row, uint32_t col) const;
    // the operating point's gmin stamps (2n words: row i's column or
    // ESP_NONE, then its slot), frozen templates only
    void gmin_stamps(std::vector<uint32_t> &out) const;
};

bool is_tridiag(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx);

// Host copy of a template's workgroup-per-lane tables (wave.h): instances
// in serial evaluation order with their staging, constant and limiter
// offsets, and the stamp tape that sums each residual row and G slot over
// staging cells in that order.
//
// Every table is also packed into one u32 blob (`blob`, word offsets in
// `at`) that a workgroup copies into LDS before its first tick.
struct InstEntry;
struct WaveBlob {
    // n_prm_lds: the leading prepared constants a workgroup copies into LDS
    // (every kind but the compiled ones, whose constants stay in HBM)
    uint32_t n, nnz, cap, n_inst, n_stage, n_prm, n_lim, tri, n_tape, n_prm_lds;
    uint64_t diag_fill;
    uint32_t col_ptr, row_idx, q, tri_pos, diag, cur_row, inst, nodes, slots, node_off, slot_off, rptr, ridx, gptr,
        gidx, qptr, qidx, cptr, cidx, words;
};
struct WaveHost {
    std::vector<uint8_t> inst_bytes;  // InstEntry array
    std::vector<uint32_t> nodes_flat, slots_flat, node_off, slot_off, rptr, ridx, gptr, gidx, qptr, qidx, cptr, cidx;
    uint32_t n_inst = 0, n_stage = 0, n_prm = 0, n_lim = 0, n_tape = 0, n_prm_lds = 0;
    WaveBlob at{};
    std::vector<uint32_t> blob;
};
void build_wave(const Template &T, uint32_t cap, WaveHost &H);

// ------------------------------------------------- a batch's lane views

constexpr uint32_t G_INITFIX = 2u;

struct Newton {
    double reltol, abstol, vntol, residual_tol, gmin;
    uint32_t max_iter, init_fix;
};

// One lane's homotopy for one static Newton of the operating point's
// stepping rungs (src/analysis/dc/op.zig:185-330): the diagonal gmin
// shunt, the gmin the compiled kinds read as $simparam("gmin") (the
// reference's setHomotopy, Circuit.zig:1394-1405), the independent
// sources' factor ("sourceScaleFactor") and whether the lane runs, cold
// (INITFIX) or warm.
struct LaneHom {
    double gshunt, gmodel, srcfact;
    uint32_t flags, pad;
};
constexpr uint32_t HOM_ACTIVE = 1u, HOM_INITFIX = 2u, HOM_PREDICT = 4u;
static_assert(sizeof(LaneHom) == 32, "LaneHom is four 8-byte words");

// A homotopy launch: per-lane controls (null: none, every lane runs as
// the Newton options say), the gmin stamps (2n words: row i's column or
// ESP_NONE, then its slot; null: every diagonal), the words of a lane's
// prepared constants that hold the compiled kinds' $simparam("gmin"), and
// for HOM_PREDICT lanes the previous point (node-major like x), against
// which a `.dc` point's predicted start is limited before its Newton
// (src/analysis/dc/dc.zig:478-483 prepareWarm).
// A mixed-precision launch (FP32 factor and solves with FP64 refinement,
// wave_mixed.h): the FP32 solves allowed per tick, and per lane two words
// out: the FP32 solves taken and the ticks handed back to FP64.
struct MixedCtl {
    uint32_t max_refine;
    uint32_t *out;
};

struct Homotopy {
    const LaneHom *lane;
    const uint32_t *stamp;
    const uint32_t *gslot;
    uint32_t n_gslot;
    const double *x_prev;
};

// Everything a lane reads or writes, as lane-major planes.
struct Batch {
    uint32_t N;
    Pattern P;
    const uint32_t *diag;
    const uint8_t *cur_row;
    int n_order;
    int order[K_NKIND];
    uint32_t count[K_NKIND];
    uint32_t kslots[K_NKIND];
    const uint32_t *nodes[K_NKIND];
    const uint32_t *slots[K_NKIND];
    const double *prm[K_NKIND];  // [(inst * PK + k) * N + l]
    double *lim[K_NKIND];        // [(inst * NU + u) * N + l]
    double *x, *x_old, *dx, *g, *rhs;
    uint32_t *status, *iters;
    LaneLUBuf lu;
};

struct PV {
    const double *p;
    size_t s;
    ESP_HD double operator[](int k) const { return p[(size_t)k * s]; }
};

// analysis/Circuit.zig evalNewton: clear g and rhs, every batch in proto
// order, the limiting correction F += J (x - x_lim), then groundStamp
// (:861-864). Ground rows and slots land in the trash row n and slot nnz.
ESP_HD void eval_planes(const Batch &B, uint32_t l, bool lim_active)
{
    const size_t N = B.N;
    const uint32_t n = B.P.n, nnz = B.P.nnz;
    LV<double> x = lane_view(B.x, N, l), g = lane_view(B.g, N, l), rhs = lane_view(B.rhs, N, l);
    for (uint32_t p = 0; p <= nnz; p++) g[p] = 0.0;
    for (uint32_t i = 0; i <= n; i++) rhs[i] = 0.0;
    for (int t = 0; t < B.n_order; t++) {
        const int kind = B.order[t];
        if (kind_is_va(kind)) continue;  // the workgroup-per-lane kernels only
        const int nu = KIND_NU[kind], pk = KIND_PK[kind];
        const unsigned limw = KIND_LIMW[kind];
        for (uint32_t id = 0; id < B.count[kind]; id++) {
            const uint32_t *nd = B.nodes[kind] + (size_t)id * nu;
            const uint32_t *sl = B.slots[kind] + (size_t)id * B.kslots[kind];
            PV p{B.prm[kind] + (size_t)id * pk * N + l, N};
            double lx[4], xl[4];
            for (int u = 0; u < nu; u++) xl[u] = lx[u] = x[nd[u]];
            const bool use_lim = limw != 0 && lim_active;
            if (use_lim) {
                const double *lm = B.lim[kind] + (size_t)id * nu * N + l;
                for (int u = 0; u < nu; u++)
                    if (limw >> u & 1) lx[u] = lm[(size_t)u * N];
            }
            DevOut o;
            eval_kind(kind, p, lx, o);
            if (use_lim) {
                double corr[4];
                bool live = false;
                for (int u = 0; u < nu; u++) {
                    corr[u] = xl[u] - lx[u];
                    live = live || corr[u] != 0;
                }
                if (live)
                    for (int ru = 0; ru < nu; ru++) {
                        double acc = 0;
                        for (int u = 0; u < nu; u++) acc += o.jac[ru][u] * corr[u];
                        o.res[ru] += acc;
                    }
            }
            uint32_t k = 0;
            for (int ru = 0; ru < nu; ru++) {
                const uint32_t row = nd[ru] == 0 ? n : nd[ru];
                if (KIND_RMASK[kind] >> ru & 1) rhs[row] += o.res[ru];
                const unsigned gm = kind_gmask(kind, ru), pm = kind_pmask(kind, ru);
                for (int cu = 0; cu < nu; cu++) {
                    if (!(pm >> cu & 1)) continue;
                    if (gm >> cu & 1) g[sl[k]] += o.jac[ru][cu];
                    k++;
                }
            }
        }
    }
    g[B.diag[0]] += 1.0;
    rhs[0] += x[0];
}

// eval.zig:1137-1168 limitRange over every limiting batch: cur = x, old =
// the previous image once limiting is live, else x_old. True when an
// instance clamped.
ESP_HD bool limit_pass(const Batch &B, uint32_t l, bool lim_active)
{
    const size_t N = B.N;
    LV<double> x = lane_view(B.x, N, l), x_old = lane_view(B.x_old, N, l);
    bool flag = false;
    for (int t = 0; t < B.n_order; t++) {
        const int kind = B.order[t];
        const unsigned limw = KIND_LIMW[kind];
        // compiled kinds run in the workgroup-per-lane kernels only
        if (limw == 0 || kind_is_va(kind)) continue;
        const int nu = KIND_NU[kind], pk = KIND_PK[kind];
        for (uint32_t id = 0; id < B.count[kind]; id++) {
            const uint32_t *nd = B.nodes[kind] + (size_t)id * nu;
            PV p{B.prm[kind] + (size_t)id * pk * N + l, N};
            double *lm = B.lim[kind] + (size_t)id * nu * N + l;
            double cur[4], old[4], lim[4];
            for (int u = 0; u < nu; u++) {
                cur[u] = x[nd[u]];
                old[u] = ((limw >> u & 1) && lim_active) ? lm[(size_t)u * N] : x_old[nd[u]];
            }
            int ch = kind == K_MOS1 ? limit_mos1(p, cur, old, lim) : limit_diode(p, cur, old, lim);
            if (ch) flag = true;
            for (int u = 0; u < nu; u++)
                if (limw >> u & 1) lm[(size_t)u * N] = lim[u];
        }
    }
    return flag;
}

// The MODEINITJCT image of a cold operating point (mos1.va:293-308,
// diode vcrit), the first limited image of the solve.
ESP_HD void seed_pass(const Batch &B, uint32_t l)
{
    const size_t N = B.N;
    LV<double> x = lane_view(B.x, N, l);
    for (int t = 0; t < B.n_order; t++) {
        const int kind = B.order[t];
        const unsigned limw = KIND_LIMW[kind];
        // compiled kinds run in the workgroup-per-lane kernels only
        if (limw == 0 || kind_is_va(kind)) continue;
        const int nu = KIND_NU[kind], pk = KIND_PK[kind];
        for (uint32_t id = 0; id < B.count[kind]; id++) {
            const uint32_t *nd = B.nodes[kind] + (size_t)id * nu;
            PV p{B.prm[kind] + (size_t)id * pk * N + l, N};
            double *lm = B.lim[kind] + (size_t)id * nu * N + l;
            double cur[4], lim[4];
            for (int u = 0; u < nu; u++) cur[u] = x[nd[u]];
            if (kind == K_MOS1) seed_mos1(p, cur, lim);
            else seed_diode(p, cur, lim);
            for (int u = 0; u < nu; u++)
                if (limw >> u & 1) lm[(size_t)u * N] = lim[u];
        }
    }
}

// The Newton gates of src/solver/kernels/lane.zig:161-214 over lane views.
ESP_HD double update_norm(uint32_t n, LV<double> x, LV<double> dx, LV<double> x_old, const uint8_t *cur_row, double reltol,
                          double abstol, double vntol)
{
    int64_t worst = 0;
    for (uint32_t i = 0; i < n; i++) {
        double xo = x[i], d = dx[i], xn = xo + d;
        x_old[i] = xo;
        x[i] = xn;
        double tol = reltol * fmax_(fabs_(xn), fabs_(xo)) + (cur_row[i] ? abstol : vntol);
        bool finite = fabs_(xn) < inf_() && fabs_(d) < inf_();
        double q = fabs_(d) / tol;
        double score = !finite ? inf_() : (q == q ? q : 0);
        int64_t b = (int64_t)bits(score);
        if (b > worst) worst = b;
    }
    return from_bits((uint64_t)worst);
}

ESP_HD bool plausible(uint32_t n, LV<double> x, double limit)
{
    for (uint32_t i = 0; i < n; i++)
        if (!(fabs_(x[i]) <= limit)) return false;
    return true;
}

ESP_HD bool residual_ok(uint32_t n, const Newton &o, LV<double> x, LV<double> rhs, LV<double> vals, const uint32_t *diag)
{
    for (uint32_t i = 0; i < n; i++) {
        double scale = fabs_(vals[diag[i]]);
        double s = scale == 0 ? 1 : scale;
        double tol = fmax_(o.residual_tol, 10.0 * s * (o.reltol * fabs_(x[i]) + o.vntol));
        if (!(fabs_(rhs[i]) <= tol)) return false;
    }
    return true;
}

ESP_HD double plausible_limit(uint32_t n, LV<double> x, LV<double> f, LV<double> vals, const uint32_t *diag)
{
    double g = 1;
    for (uint32_t i = 0; i < n; i++) g = fmax_(g, fabs_(vals[diag[i]]));
    uint64_t xm = 0, fm = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint64_t bx = bits(fabs_(x[i])), bf = bits(fabs_(f[i]));
        if (bx > xm) xm = bx;
        if (bf > fm) fm = bf;
    }
    double scale = fmin_(fmax_(from_bits(xm), from_bits(fm) * g), 1e18);
    return !(1e3 * scale > 1e9) ? 1e9 : 1e3 * scale;
}

// One lane's static Newton (converger.zig:155-319 newton with
// finalizeStep :458-526; the device ticks of lane_device.zig :316-384):
// assemble, the lane's own direct factor and solve, update and delta
// score, limiting, then the gates in the reference's order. Status bits:
// 1 converged, 2 singular. A converged lane publishes x_k (x = x_old).
ESP_HD void lane_newton(const Batch &B, uint32_t l, const Newton &o)
{
    const size_t N = B.N;
    const uint32_t n = B.P.n;
    LV<double> x = lane_view(B.x, N, l), x_old = lane_view(B.x_old, N, l), dx = lane_view(B.dx, N, l);
    LV<double> g = lane_view(B.g, N, l), rhs = lane_view(B.rhs, N, l);
    LaneLU s = lane_lu(B.lu, N, l, n);
    uint32_t bits_ = o.init_fix ? G_INITFIX : 0;
    bool lim_active = false;
    if (o.init_fix) {
        seed_pass(B, l);
        lim_active = true;
    }
    double limit = 1e9;
    for (uint32_t i = 0; i < n; i++) x_old[i] = x[i];
    B.status[l] = 0;
    B.iters[l] = o.max_iter;
    for (uint32_t tick = 0; tick < o.max_iter; tick++) {
        eval_planes(B, l, lim_active);
        if (o.gmin > 0) {
            for (uint32_t i = 0; i < n; i++) g[B.diag[i]] += o.gmin;
            for (uint32_t i = 0; i < n; i++) rhs[i] += o.gmin * x[i];
        }
        if (tick == 0) limit = plausible_limit(n, x, rhs, g, B.diag);
        if (direct_factor(B.P, s, g) != 0) {
            B.status[l] = 2;
            B.iters[l] = 0;
            return;
        }
        direct_solve_neg(B.P, s, rhs, dx);
        double scaled = update_norm(n, x, dx, x_old, B.cur_row, o.reltol, o.abstol, o.vntol);
        bool limited = limit_pass(B, l, lim_active);
        lim_active = true;
        bool ok = !limited && tick != 0 && !(scaled >= 1.0);
        if (ok) ok = plausible(n, x, limit);
        if (ok) ok = residual_ok(n, o, x, rhs, g, B.diag);
        if (ok && (bits_ & G_INITFIX)) {
            bits_ &= ~G_INITFIX;
        } else if (ok) {
            for (uint32_t i = 0; i < n; i++) x[i] = x_old[i];
            B.status[l] = 1;
            B.iters[l] = tick + 1;
            return;
        }
    }
}

}  // namespace esp
