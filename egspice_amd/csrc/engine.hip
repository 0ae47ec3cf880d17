// The lane engine: circuit templates, lane batches resident on one MI355X,
// the gfx950 kernels, and the C ABI of include/espice_lanes.h.
//
// Kernels:
//   newton_wave_kernel    the whole static Newton of a lane in one launch,
//                         one 64-thread workgroup per lane with the lane's
//                         state and the template's tables in LDS (wave.h):
//                         device evaluation into staging cells, the ordered
//                         stamp reduction, the lane's own direct factor and
//                         solve (wave-parallel for up to 64 unknowns,
//                         wave_lu.h), update, limiting and the gates, per
//                         tick, until the lane converges, goes singular or
//                         runs out of ticks. The reference runs the same tick
//                         as two captured graphs of ~10 kernels with a host
//                         sync per tick (src/analysis/gpu/lanes.zig:
//                         1431-1601); here the tick loop lives on the device,
//                         so a solve is one launch and one sync.
//   newton_op_kernel      the same Newton one thread per lane over lane-major
//                         global planes, for templates whose lane block does
//                         not fit in LDS.
//   tab_refactor_kernel   the reference's `arp_lane_refactor` + `arp_lane_solve`
//                         over a shared lead tape (lane_device.zig:101,170).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <strings.h>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "../../include/espice_lanes.h"
#include "engine.h"
#include "wave.h"
#include "wave_lu.h"
#include "wave_newton.h"
#include "wave_tran.h"
#include "mlp.h"

using namespace esp;

namespace {

thread_local std::string g_error;

espice_status fail(espice_status s, const char *name)
{
    g_error = name;
    return s;
}

// A HIP error names itself and the engine line that met it.
espice_status fail_hip(hipError_t e, int line)
{
    g_error = std::string(hipGetErrorName(e)) + " (engine.hip:" + std::to_string(line) + ")";
    return ESPICE_FAILED;
}

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return fail_hip(e_, __LINE__); \
    } while (0)

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count)
    {
        n = count ? count : 1;
        return hipMalloc(reinterpret_cast<void **>(&p), n * sizeof(T));
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// Page-locked host memory mapped for the GPU: the workgroup-per-lane
// kernels write their per-lane results straight into it, so a solve ends
// with one synchronisation and no copies.
template <typename T>
struct HostBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count)
    {
        n = count ? count : 1;
        return hipHostMalloc(reinterpret_cast<void **>(&p), n * sizeof(T), hipHostMallocMapped);
    }
    ~HostBuf()
    {
        if (p) (void)hipHostFree(p);
    }
};

}  // namespace

void esp::set_error(const char *name) { g_error = name; }

struct espice_lanes_template {
    Template t;
};

// ---------------------------------------------------------------- kernels

__global__ __launch_bounds__(64) void newton_op_kernel(Batch B, Newton o)
{
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= B.N) return;
    lane_newton(B, l, o);
}

// Workgroup-per-lane static Newton (wave.h): block l owns lane l, its
// state in dynamic LDS laid out by `lay` after the template's tables. With
// DENSE the lane's direct solver replays on the whole wavefront
// (wave_lu.h, patterns up to 64 unknowns); otherwise thread 0 runs it.
// With a homotopy (hom.lane), a lane whose control is inactive returns at
// once; the others take their gmin shunt, cold or warm start and source
// factor from it and first write their model gmin into their own
// prepared constants.
// MIXED: every tick's linear step in FP32 with FP64 refinement, its arrays
// after the lane block (mixed_words), tallies to mix.out.
template <bool DENSE, int VA, bool MIXED = false>
__global__ __launch_bounds__(64) void newton_wave_kernel(WaveBlob B, const uint32_t *blob, WaveLayout lay,
                                                         double *prm_lanes, double *x_glob, uint32_t N, Newton o_in,
                                                         uint32_t *status, uint32_t *iters, const uint32_t *pred,
                                                         uint32_t *pred_hit, uint32_t *h_status, uint32_t *h_iters,
                                                         double *h_x, double *lu_g, Homotopy hom, MixedCtl mix)
{
    static_assert(DENSE || !MIXED, "the mixed step takes a dense lane (n <= 64)");
    extern __shared__ double lds[];
    const uint32_t l = blockIdx.x;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t n = B.n;
    Newton o = o_in;
    double srcfact = 1.0;
    if (hom.lane != nullptr) {
        const LaneHom h = hom.lane[l];
        if (!(h.flags & HOM_ACTIVE)) return;
        o.gmin = h.gshunt;
        o.init_fix = (h.flags & HOM_INITFIX) ? 1u : 0u;
        srcfact = h.srcfact;
        for (uint32_t i = tid; i < hom.n_gslot; i += nt) prm_lanes[(size_t)l * B.n_prm + hom.gslot[i]] = h.gmodel;
        __threadfence_block();
        __syncthreads();
    }
    // the template's tables first (B.words is even; the lean layout reads
    // them in global memory), then the lane block
    uint32_t *tl = reinterpret_cast<uint32_t *>(lds);
    if (!lay.lu_global)
        for (uint32_t i = tid; i < B.words; i += nt) tl[i] = blob[i];
    double *block = lds + table_lds_words(B.words, lay) / 2;
    const WaveTemplate W = wave_template(B, lay.lu_global ? blob : tl);
    WaveLane s = wave_lane(W, lay, block);
    if (hom.lane != nullptr) {
        s.gstamp = hom.stamp;
        s.srcfact = srcfact;
    }
    if (lay.lu_global) wave_bind_lu_scratch(s, W, lu_g + (size_t)l * lane_scratch_words(W.cap, W.n_tape));
    DenseLU D = dense_lu(block, lay, n, pred, pred_hit ? pred_hit + l : nullptr);
    if (tid == 0 && pred_hit) pred_hit[l] = 0;  // this launch's record starts empty
    // the lane's first factor will try the prediction: its tape's loads go
    // out with the template's
    if (DENSE && pred != nullptr) {
        pred_stage(W.P, s.lu, D);
        D.staged = true;
    }
    double *prm_lds = block + lay.prm;
    for (uint32_t i = tid; i < W.n_prm_lds; i += nt) prm_lds[i] = prm_lanes[(size_t)l * W.n_prm + i];
    s.gprm = prm_lanes + (size_t)l * W.n_prm;
    for (uint32_t i = tid; i < n; i += nt) {
        double v = x_glob[(size_t)i * N + l];
        s.x[i] = v;
        s.x_old[i] = v;
    }
    __syncthreads();
    if (tid == 0) wave_fresh_solver(W, s);
    __syncthreads();
    bool lim_active = false;
    if (o.init_fix) {
        for (uint32_t e = tid; e < W.n_inst; e += nt) wave_seed_inst<VA>(W, s, e);
        lim_active = true;
    } else if (hom.lane != nullptr && (hom.lane[l].flags & HOM_PREDICT)) {
        // the predicted start limited against the previous point, the
        // Newton then starting from the limiters' images
        for (uint32_t i = tid; i < n; i += nt) s.x_old[i] = hom.x_prev[(size_t)i * N + l];
        __syncthreads();
        for (uint32_t e = tid; e < W.n_inst; e += nt) (void)wave_limit_inst<VA>(W, s, e, false);
        __syncthreads();
        for (uint32_t i = tid; i < n; i += nt) s.x_old[i] = s.x[i];
        lim_active = true;
    }
    const ReduceCache rc = wave_reduce_cache(W, tid);
    uint32_t st = 0, it = o.max_iter;
    __syncthreads();
    if (DENSE) {
        const WaveHook off{};
        MixedRun mx{};
        if constexpr (MIXED) {
            mx.m = mixed_lu(block + lay.words, n);
            mx.max_refine = mix.max_refine;
            mixed_rows(W.P, mx.m);
        }
        const int r = wave_newton<false, false, VA, MIXED>(W, s, D, rc, o, off, o.init_fix != 0, lim_active, 0.0, it,
                                                           &mx);
        st = r == NEWTON_CONVERGED ? 1u : r == NEWTON_SINGULAR ? 2u : 0u;
        if (MIXED && tid == 0) {
            mix.out[2 * (size_t)l] = mx.solves;
            mix.out[2 * (size_t)l + 1] = mx.fallbacks;
        }
    } else {
        uint32_t bits = o.init_fix ? G_INITFIX : 0;
        double limit = 1e9, scaled = 0;
        for (uint32_t tick = 0; tick < o.max_iter; tick++) {
            for (uint32_t e = tid; e < W.n_inst; e += nt) wave_eval_inst<false, VA>(W, s, e, lim_active);
            __syncthreads();
            wave_reduce(W, s, rc, tid, nt);
            __syncthreads();
            if (tid == 0) s.flags[0] = wave_solve(W, s, o, tick, limit, scaled) ? 0u : 1u;
            __syncthreads();
            if (s.flags[0]) {
                st = 2;
                it = 0;
                break;
            }
            int lim_any = 0;
            for (uint32_t e = tid; e < W.n_inst; e += nt) lim_any |= wave_limit_inst<VA>(W, s, e, lim_active) ? 1 : 0;
            lim_any = __syncthreads_or(lim_any);
            lim_active = true;
            if (tid == 0) {
                bool ok = wave_gates(W, s, o, tick, lim_any != 0, limit, scaled);
                uint32_t verdict = 0;
                if (ok && (bits & G_INITFIX)) bits &= ~G_INITFIX;
                else if (ok) verdict = 1;
                s.flags[1] = verdict;
            }
            __syncthreads();
            if (s.flags[1]) {
                for (uint32_t i = tid; i < n; i += nt) s.x[i] = s.x_old[i];
                st = 1;
                it = tick + 1;
                break;
            }
            __syncthreads();
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < n; i += nt) x_glob[(size_t)i * N + l] = s.x[i];
    if (h_x)
        for (uint32_t i = tid; i < n; i += nt) h_x[(size_t)i * N + l] = s.x[i];
    if (tid == 0) {
        status[l] = st;
        iters[l] = it;
        h_status[l] = st;
        h_iters[l] = it;
    }
}

// The template's predicted first factor: lane 0's matrix at a cold start
// (x = 0, the limiters' seeds) factored by the serial full factor, its
// symbolic tape written to `pred` (PredTape) for every lane's first
// factor to try; PRED_OK stays 0 when that factor fails or voids a
// column. One workgroup; setup work, not part of a solve.
template <int VA>
__global__ __launch_bounds__(64) void analyze_wave_kernel(WaveBlob B, const uint32_t *blob, WaveLayout lay,
                                                          const double *prm_lanes, Newton o, uint32_t *pred,
                                                          double *lu_g)
{
    extern __shared__ double lds[];
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t n = B.n, nnz = B.nnz;
    uint32_t *tl = reinterpret_cast<uint32_t *>(lds);
    if (!lay.lu_global)
        for (uint32_t i = tid; i < B.words; i += nt) tl[i] = blob[i];
    double *block = lds + table_lds_words(B.words, lay) / 2;
    const WaveTemplate W = wave_template(B, lay.lu_global ? blob : tl);
    WaveLane s = wave_lane(W, lay, block);
    if (lay.lu_global) wave_bind_lu_scratch(s, W, lu_g);
    const DenseLU D = dense_lu(block, lay, n);
    double *prm_lds = block + lay.prm;
    for (uint32_t i = tid; i < W.n_prm_lds; i += nt) prm_lds[i] = prm_lanes[i];
    s.gprm = prm_lanes;
    for (uint32_t i = tid; i < n; i += nt) {
        s.x[i] = 0.0;
        s.x_old[i] = 0.0;
    }
    __syncthreads();
    if (tid == 0) wave_fresh_solver(W, s);
    __syncthreads();
    for (uint32_t e = tid; e < W.n_inst; e += nt) wave_seed_inst<VA>(W, s, e);
    const ReduceCache rc = wave_reduce_cache(W, tid);
    __syncthreads();
    for (uint32_t e = tid; e < W.n_inst; e += nt) wave_eval_inst<false, VA>(W, s, e, true);
    __syncthreads();
    wave_reduce(W, s, rc, tid, nt);
    __syncthreads();
    if (tid == 0) {
        double limit = 1e9;
        wave_assemble(W, s, o, 0, limit);
        bool ok = lu_factor(W.P, s.lu, s.g) == 0 && s.lu.cnt[C_NVOID] == 0;
        for (uint32_t i = 0; i < n; i++) ok = ok && s.lu.void_col[i] == 0;
        s.flags[2] = ok ? 1u : 0u;
        pred[PRED_OK] = 0;
    }
    __syncthreads();
    if (!s.flags[2]) return;
    dense_build(W.P, s.lu, D);
    const PredTape T = pred_tape(n, nnz);
    const uint32_t nl = s.lu.cnt[C_NL], nu = s.lu.cnt[C_NU];
    for (uint32_t i = tid; i < n; i += nt) {
        pred[T.pinv + i] = s.lu.pinv[i];
        pred[T.porig + i] = D.porig[i];
        pred[T.lmask + 2 * i] = (uint32_t)D.lmask[i];
        pred[T.lmask + 2 * i + 1] = (uint32_t)(D.lmask[i] >> 32);
        pred[T.umask + 2 * i] = (uint32_t)D.umask[i];
        pred[T.umask + 2 * i + 1] = (uint32_t)(D.umask[i] >> 32);
    }
    for (uint32_t i = tid; i <= n; i += nt) {
        pred[T.lp + i] = s.lu.lp[i];
        pred[T.up + i] = s.lu.up[i];
    }
    for (uint32_t i = tid; i < nl; i += nt) pred[T.li + i] = s.lu.li[i];
    for (uint32_t i = tid; i < nu; i += nt) pred[T.ui + i] = s.lu.ui[i];
    for (uint32_t i = tid; i < nnz; i += nt) pred[T.prow + i] = s.lu.prow[i];
    for (uint32_t i = tid; i < n * n; i += nt) pred[T.amap + i] = amap_slot(D, i);
    if (tid == 0) {
        pred[PRED_NL] = nl;
        pred[PRED_NU] = nu;
        pred[PRED_OK] = 1;
    }
}

// One lane's transient per workgroup (wave_tran.h): the operating point at
// t = 0, then steps until t_stop, each accepted row fed to the .meas
// passes and, up to `to.record_cap` rows, recorded. The lane's final
// state goes back to x_glob.
template <int VA, bool MIXED = false>
__global__ __launch_bounds__(64) void tran_wave_kernel(WaveBlob B, const uint32_t *blob, WaveLayout lay,
                                                       const double *prm_lanes, double *x_glob, uint32_t N,
                                                       Newton op_opts, Newton step_opts, TranOpts to,
                                                       const MeasPass *passes, uint32_t n_pass, TranOut *outs,
                                                       MeasOut *meas_out, double *rec_t, double *rec_x,
                                                       const uint32_t *pred, double *lu_g, MixedCtl mix)
{
    extern __shared__ double lds[];
    const uint32_t l = blockIdx.x;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t n = B.n, nnz = B.nnz;
    uint32_t *tl = reinterpret_cast<uint32_t *>(lds);
    if (!lay.lu_global)
        for (uint32_t i = tid; i < B.words; i += nt) tl[i] = blob[i];
    double *block = lds + table_lds_words(B.words, lay) / 2;
    const WaveTemplate W = wave_template(B, lay.lu_global ? blob : tl);
    WaveLane s = wave_lane(W, lay, block);
    double *const lane_scr = lu_g + (size_t)l * lane_scratch_words(W.cap, W.n_tape);
    if (lay.lu_global) wave_bind_lu_scratch(s, W, lane_scr);
    const DenseLU D = dense_lu(block, lay, n, pred);
    // MIXED: the Newton ticks' linear steps in FP32 with FP64 refinement,
    // their arrays after the lane block
    MixedRun mx{};
    if constexpr (MIXED) {
        mx.m = mixed_lu(block + lay.words, n);
        mx.max_refine = mix.max_refine;
        mixed_rows(W.P, mx.m);
    }
    const uint32_t n_tape = W.n_tape;
    double *const a = s.cc + (nnz + 1);
    double *const ring = s.tape + n_tape;
    double *const i_prev = ring + 4 * (size_t)n;
    // the charge sites' history: in the lane's HBM scratch in the lean
    // layout (read and written once per step), else in the block
    double *const tring = lay.lu_global ? lane_scr + 3 * (size_t)W.cap : i_prev + n;
    double *const ti_prev = tring + 4 * (size_t)n_tape;
    double *const cur = lay.lu_global ? i_prev + n : ti_prev + n_tape;
    double *const prev = cur + n;
    double *const scal = prev + n;
    double *const bpc = scal + TRAN_SCALARS;
    MeasState *const ms = reinterpret_cast<MeasState *>(bpc + 2 * (size_t)W.n_inst);
    double *prm_lds = block + lay.prm;
    for (uint32_t i = tid; i < W.n_prm_lds; i += nt) prm_lds[i] = prm_lanes[(size_t)l * W.n_prm + i];
    s.gprm = prm_lanes + (size_t)l * W.n_prm;
    for (uint32_t i = tid; i < n; i += nt) {
        // a cold operating point starts at 0; a warm one at the lane's x
        s.x[i] = to.op_warm ? x_glob[(size_t)i * N + l] : 0.0;
        i_prev[i] = 0.0;
    }
    for (uint32_t j = tid; j < n_tape; j += nt) ti_prev[j] = 0.0;
    // the charge history starts at zero: a template without charge sites
    // never seeds or rotates it, and its companion terms must read 0, not
    // whatever an earlier launch left in LDS
    for (uint32_t i = tid; i < 4 * n; i += nt) ring[i] = 0.0;
    for (uint32_t j = tid; j < 4 * n_tape; j += nt) tring[j] = 0.0;
    __syncthreads();
    // the .tran card's PULSE, SIN and EXP defaults (frontend/builder.zig:845-877)
    for (uint32_t e = tid; e < W.n_inst; e += nt) {
        const InstEntry ie = W.inst[e];
        bpc[2 * e] = inf_();
        bpc[2 * e + 1] = inf_();
        if (ie.kind != K_VSRC && ie.kind != K_ISRC) continue;
        double *p = prm_lds + ie.prm;
        if (p[SRCP_WAVE] == 5.0) {
            if (p[SRCP_FFM] == -1.0) p[SRCP_FFM] = 5.0 / to.t_stop;
            if (p[SRCP_FFC] == 0.0) p[SRCP_FFC] = 500.0 / to.t_stop;
        }
        if (p[SRCP_WAVE] == 2.0 && p[SRCP_SFREQ] == 0) p[SRCP_SFREQ] = 1.0 / to.t_stop;
        if (p[SRCP_WAVE] == 3.0) {
            if (p[SRCP_ETD1] == 0) p[SRCP_ETD1] = to.dt_init;
            if (p[SRCP_ETAU1] == 0) p[SRCP_ETAU1] = to.dt_init;
            if (p[SRCP_ETD2] == 0) p[SRCP_ETD2] = p[SRCP_ETD1] + to.dt_init;
            if (p[SRCP_ETAU2] == 0) p[SRCP_ETAU2] = to.dt_init;
        }
        if (p[SRCP_WAVE] != 1.0) {
            p[SRCP_TD] = 1e30;
        } else {
            if (p[SRCP_TR] <= 0) p[SRCP_TR] = to.dt_init;
            if (p[SRCP_TF] <= 0) p[SRCP_TF] = to.dt_init;
            if (p[SRCP_PW] <= 0) p[SRCP_PW] = to.t_stop;
            if (p[SRCP_PER] <= 0) p[SRCP_PER] = to.t_stop;
        }
    }
    if (tid == 0) {
        wave_fresh_solver(W, s);
        for (uint32_t k = 0; k < n_pass; k++) meas_init(passes[k], ms[k]);
    }
    __syncthreads();
    if (!to.op_warm)
        for (uint32_t e = tid; e < W.n_inst; e += nt) wave_seed_inst<VA>(W, s, e);
    const ReduceCache rc = wave_reduce_cache(W, tid);
    __syncthreads();

    TranStep st{};
    uint32_t it = 0;
    const WaveHook off{};
    // cold: MODEINITJCT seeds and INITFIX; warm: Newton from x, as the
    // operating point's warm start (src/analysis/dc/op.zig:588, 671)
    const bool cold = to.op_warm == 0;
    const int op = wave_newton<true, false, VA, MIXED>(W, s, D, rc, op_opts, off, cold, cold, 0.0, it, &mx);
    if (op != NEWTON_CONVERGED) {
        if (tid == 0) {
            st.out.status = op == NEWTON_SINGULAR ? 3u : 2u;
            outs[l] = st.out;
            const double nan = __builtin_nan("");
            for (uint32_t k = 0; k < n_pass; k++) meas_out[(size_t)l * n_pass + k] = MeasOut{nan, nan, nan};
        }
        return;
    }
    for (uint32_t i = tid; i < n; i += nt) {
        cur[i] = s.x[i];
        prev[i] = s.x[i];
    }
    const bool has_charge = n_tape > 0;
    if (has_charge) {
        // q(op) seeds a flat history
        for (uint32_t e = tid; e < W.n_inst; e += nt) wave_eval_inst<true, VA, true>(W, s, e, false, 0.0);
        __syncthreads();
        for (uint32_t e = tid; e < n; e += nt) wave_reduce_charge(W, s, e);
        __syncthreads();
        for (uint32_t k = 1; k < 4; k++) {
            for (uint32_t i = tid; i < n; i += nt) ring[k * (size_t)n + i] = s.q[i];
            for (uint32_t j = tid; j < n_tape; j += nt) tring[k * (size_t)n_tape + j] = s.tape[j];
        }
    }
    const size_t cap = to.record_cap;
    uint32_t points = 1;
    if (cap > 0)
        for (uint32_t i = tid; i < n; i += nt) rec_x[((size_t)l * cap) * n + i] = cur[i];
    __syncthreads();
    if (tid == 0) {
        if (cap > 0) rec_t[(size_t)l * cap] = 0.0;
        for (uint32_t k = 0; k < n_pass; k++) {
            if (passes[k].kind == 4) meas_feed_find(passes[k], ms[k], ms[k - 1], 0.0, cur[passes[k].row]);
            else meas_feed(passes[k], ms[k], 0.0, cur[passes[k].row]);
        }
        st.delmin = 1e-11 * to.dt_max;
        st.min_break = 10.0 * st.delmin;
        st.bp_lo = inf_();
        st.bp_hi = -inf_();
        double dt = zmin(zmin(to.dt_init, to.t_stop / 100.0) / 10.0, to.dt_max);
        const double bp0 = tran_next_bp(W, prm_lds, bpc, st, 0.0, st.min_break);
        if (bp0 != inf_()) dt = zmin(dt, 0.1 * bp0);
        st.dt = dt / 10.0;
        for (int k = 0; k < 6; k++) st.dt_old[k] = to.dt_max;
        st.order = st.run_order = 1;
        st.bp_next = inf_();
        st.bp_save_dt = to.t_stop / 50.0;
        for (uint32_t k = 0; k < 4; k++) st.perm[k] = k;
        st.out.points = 1;
    }
    for (;;) {
        // Sim.begin
        if (tid == 0) {
            const bool go = !st.stopped && st.t < to.t_stop && !ulps_close(st.t, to.t_stop, 100) && st.steps < to.max_steps;
            scal[0] = go ? 1.0 : 0.0;
            if (go) {
                st.dt_att = st.dt;
                const int method = st.order == 1 ? 0 : (int)to.method;
                const TranCoeffs c = tran_coeffs(method, st.dt, st.dt_old[0], to.xmu);
                scal[1] = st.t + st.dt;
                scal[2] = st.dt / st.dt_old[0];
                scal[3] = (double)method;
                scal[4] = c.ag0;
                scal[5] = c.ag1;
                scal[6] = c.ag2;
                for (uint32_t k = 0; k < 4; k++) scal[8 + k] = (double)st.perm[k];
                scal[12] = st.dt;
                scal[13] = st.dt_old[0];
                scal[14] = st.dt_old[1];
            }
        }
        __syncthreads();
        if (scal[0] == 0.0) break;
        const double t_dev = scal[1], xfact = scal[2];
        const int method = (int)scal[3];
        const uint32_t p0 = (uint32_t)scal[8], p1 = (uint32_t)scal[9], p2 = (uint32_t)scal[10];
        for (uint32_t i = tid; i < n; i += nt) {
            s.x[i] = cur[i] + xfact * (cur[i] - prev[i]);
            s.x_old[i] = cur[i];
        }
        __syncthreads();
        // MODEINITPRED: the predicted iterate limited against the last
        // accepted point, the limiters cleared by the previous solve
        // (tran.zig:1185-1189, eval.zig:2430-2434); Newton then starts with
        // limiting active
        for (uint32_t e = tid; e < W.n_inst; e += nt) (void)wave_limit_inst<VA>(W, s, e, false);
        __syncthreads();
        WaveHook h{method, scal[4], scal[5], scal[6], ring + p1 * (size_t)n, ring + p2 * (size_t)n, i_prev, a};
        const int r = wave_newton<true, true, VA, MIXED>(W, s, D, rc, step_opts, h, false, true, t_dev, it, &mx);
        if (r == NEWTON_CONVERGED && has_charge) {
            // the charge re-read at the solution, unlimited (tran.zig:859-864,
            // eval.zig:1086-1108)
            for (uint32_t e = tid; e < W.n_inst; e += nt) wave_eval_inst<true, VA, true>(W, s, e, false, t_dev);
            __syncthreads();
            for (uint32_t e = tid; e < n; e += nt) wave_reduce_charge(W, s, e);
            __syncthreads();
            for (uint32_t i = tid; i < n; i += nt) ring[p0 * (size_t)n + i] = s.q[i];
            for (uint32_t j = tid; j < n_tape; j += nt) tring[p0 * (size_t)n_tape + j] = s.tape[j];
        }
        __syncthreads();
        // the step's two LTE bounds (its own method and, for a promotion,
        // the run's), over the charge sites across the wavefront
        double del_w = 0.0, trial_w = 0.0;
        if (r == NEWTON_CONVERGED && has_charge) {
            const TranCoeffs c{method, scal[4], scal[5], scal[6]};
            const LteIn e{scal[12], scal[13], scal[14], to.reltol, to.abstol, to.chgtol, to.trtol};
            const uint32_t q3 = (uint32_t)scal[11];
            const double *t0p = tring + p0 * (size_t)n_tape, *t1p = tring + p1 * (size_t)n_tape;
            const double *t2p = tring + p2 * (size_t)n_tape, *t3p = tring + q3 * (size_t)n_tape;
            del_w = wave_step_bound(method, method, t0p, t1p, t2p, t3p, ti_prev, n_tape, c, e);
            trial_w = wave_step_bound((int)to.method, method, t0p, t1p, t2p, t3p, ti_prev, n_tape, c, e);
        }
        // Sim.end up to the commit
        if (tid == 0) {
            const TranCoeffs c{method, scal[4], scal[5], scal[6]};
            const double dt = st.dt;
            bool accept = false;
            st.out.attempts++;
            st.out.nr_iters += r == NEWTON_SINGULAR ? 0u : it;
            if (r != NEWTON_CONVERGED) {
                st.out.rej_newton++;
                st.order = 1;
                st.dt = dt / 8.0;
                tran_floor_retry(to, st);
            } else {
                double dt_next = st.steps == 0 ? dt : zmin(dt * 2.0, to.dt_max);
                accept = true;
                if (has_charge && st.steps > 0) {
                    const double del = del_w;
                    if (!(del >= 0.9 * dt)) {
                        st.out.rej_lte++;
                        if (del != del) {
                            st.order = 1;
                            st.dt = dt / 8.0;
                        } else {
                            st.dt = del;
                        }
                        tran_floor_retry(to, st);
                        accept = false;
                    } else {
                        dt_next = zmin(zmin(sel_max(del, to.dt_min), 2.0 * dt), to.dt_max);
                    }
                }
                if (accept && has_charge) {
                    st.order_run = st.order == st.run_order ? st.order_run + 1 : 1;
                    st.run_order = st.order;
                    if (st.steps > 0 && st.order == 1) {
                        const double trial_del = trial_w;
                        if (trial_del == trial_del) {
                            const double nd2 = zmin(2.0 * dt, trial_del);
                            if (nd2 > 1.05 * dt) st.order = 2;
                            dt_next = zmin(sel_max(nd2, to.dt_min), to.dt_max);
                        }
                    }
                }
                st.dt_next = dt_next;
            }
            scal[7] = accept ? 1.0 : 0.0;
        }
        __syncthreads();
        if (scal[7] == 0.0) continue;
        // advanceCurrent on the pre-rotation rings, then cur <- trial
        if (has_charge) {
            const double ag0 = scal[4], ag1 = scal[5], ag2 = scal[6];
            for (uint32_t i = tid; i < n; i += nt) {
                const double q0v = ring[p0 * (size_t)n + i], q1v = ring[p1 * (size_t)n + i];
                const double d = ag0 * (q0v - q1v);
                i_prev[i] = method == 1 ? d - ag1 * i_prev[i] : method == 2 ? d - ag2 * (q1v - ring[p2 * (size_t)n + i]) : d;
            }
            for (uint32_t j = tid; j < n_tape; j += nt) {
                const double q0v = tring[p0 * (size_t)n_tape + j], q1v = tring[p1 * (size_t)n_tape + j];
                const double d = ag0 * (q0v - q1v);
                ti_prev[j] = method == 1 ? d - ag1 * ti_prev[j]
                             : method == 2 ? d - ag2 * (q1v - tring[p2 * (size_t)n_tape + j])
                                           : d;
            }
        }
        for (uint32_t i = tid; i < n; i += nt) {
            prev[i] = cur[i];
            cur[i] = s.x[i];
        }
        if (points < cap)
            for (uint32_t i = tid; i < n; i += nt) rec_x[((size_t)l * cap + points) * n + i] = s.x[i];
        points++;
        __syncthreads();
        // Sim.end after the commit: rotation, landing, the row, next dt
        if (tid == 0) {
            if (has_charge) {
                const uint32_t last = st.perm[3];
                for (int k = 3; k > 0; k--) st.perm[k] = st.perm[k - 1];
                st.perm[0] = last;
            }
            for (int k = 5; k > 0; k--) st.dt_old[k] = st.dt_old[k - 1];
            st.dt_old[0] = st.dt;
            st.t = st.t + st.dt;
            st.steps++;
            const double t_new = st.t;
            double dt_next = st.dt_next;
            if (st.bp_next != inf_()) {
                const double bp = st.bp_next;
                const bool landed = st.bp_clamped ? fabs_(t_new - bp) <= st.min_break
                                                  : (bp - t_new <= st.delmin || ulps_close(t_new, bp, 100));
                if (landed) {
                    st.out.bp_landings++;
                    st.order = 1;
                    double shrink = st.bp_save_dt;
                    const double nb = tran_next_bp(W, prm_lds, bpc, st, t_new, t_new + st.min_break);
                    if (nb != inf_()) shrink = zmin(shrink, nb - t_new);
                    dt_next = zmin(dt_next, 0.1 * shrink);
                }
                st.bp_next = inf_();
            }
            if (st.out.points < cap) rec_t[(size_t)l * cap + st.out.points] = t_new;
            st.out.points++;
            bool open = false, bounded_open = false;
            for (uint32_t k = 0; k < n_pass; k++) {
                if (passes[k].kind == 4) meas_feed_find(passes[k], ms[k], ms[k - 1], t_new, cur[passes[k].row]);
                else meas_feed(passes[k], ms[k], t_new, cur[passes[k].row]);
                open = open || !ms[k].i[7];
                bounded_open = bounded_open || (!ms[k].i[7] && k < 64 && (to.abort_mask >> k & 1));
            }
            // the point's measure bound and early stop (tran.zig:1059-1064):
            // past abort_after with a bounded measure unresolved the lane
            // fails; with every measure resolved an early-stopping lane ends
            if (to.abort_after == to.abort_after && t_new > to.abort_after && bounded_open) {
                st.out.status = 4;
                st.stopped = true;
            } else if (to.early_stop && n_pass > 0 && !open) {
                st.stopped = true;
            }
            st.bp_next = tran_next_bp(W, prm_lds, bpc, st, t_new, t_new + st.min_break);
            st.bp_clamped = false;
            if (st.bp_next != inf_()) {
                const double dt_to_bp = st.bp_next - t_new;
                if (dt_to_bp < dt_next) {
                    st.bp_save_dt = dt_next;
                    dt_next = dt_to_bp;
                    st.bp_clamped = true;
                }
            }
            st.dt = dt_next;
            if (t_new + st.dt > to.t_stop) st.dt = to.t_stop - t_new;
        }
        __syncthreads();
    }
    for (uint32_t i = tid; i < n; i += nt) x_glob[(size_t)i * N + l] = cur[i];
    if (tid == 0) {
        st.out.steps = st.steps;
        st.out.fp32_solves = mx.solves;
        st.out.fp64_fallbacks = mx.fallbacks;
        outs[l] = st.out;
        for (uint32_t k = 0; k < n_pass; k++) meas_out[(size_t)l * n_pass + k] = meas_result(passes[k], ms[k]);
    }
}

__global__ void init_lane_state_kernel(uint8_t *state, uint32_t N, uint8_t tri)
{
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= N) return;
    state[(size_t)ST_FACTORED * N + l] = 0;
    state[(size_t)ST_TRI * N + l] = tri;
    state[(size_t)ST_LU_FACTORED * N + l] = 0;
}

__global__ __launch_bounds__(64) void tab_refactor_kernel(LaneTab t, const uint32_t *idx, uint32_t N, const double *a,
                                                          const double *rhs, double *dx, uint8_t *fail, double *lx,
                                                          double *ux, double *udiag, double *w, double *y)
{
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= N) return;
    LV<const double> av{a + l, N}, rv{rhs + l, N};
    LV<double> dv{dx + l, N};
    bool bad = tab_refactor(t, idx, av, LV<double>{lx + l, N}, LV<double>{ux + l, N}, LV<double>{udiag + l, N},
                            LV<double>{w + l, N});
    fail[l] = bad ? 1 : 0;
    if (!bad) tab_solve_neg(t, idx, LV<double>{lx + l, N}, LV<double>{ux + l, N}, LV<double>{udiag + l, N}, rv,
                            LV<double>{y + l, N}, dv);
}

// ------------------------------------------------------- AC lanes

// One lane's small-signal response at its operating point
// (src/analysis/ac/ac.zig:40-86, one lane of freq.Stream): the lane's G
// and C planes once, from one transient-mode evaluation at x (no limiting;
// the ground row pinned), then for each frequency (G + j w C) X = drive
// factored densely in complex FP64 with partial pivoting (largest
// |re| + |im|, the lowest row among equals) on the wavefront and solved,
// thread r owning row r of the vectors. Out: [(k * 2 + part) * n + i] * N
// + l, part 0 the real planes, 1 the imaginary; NaN where the matrix is
// singular. The complex matrix lives after the lane block (2 n * n doubles).
template <int VA>
__global__ __launch_bounds__(64) void ac_wave_kernel(WaveBlob B, const uint32_t *blob, WaveLayout lay,
                                                     const double *prm_lanes, const double *x_glob, uint32_t N,
                                                     const double *omega, uint32_t n_freq, const double *drive,
                                                     double *out, double *lu_g)
{
    extern __shared__ double lds[];
    const uint32_t l = blockIdx.x;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t n = B.n, nnz = B.nnz;
    uint32_t *tl = reinterpret_cast<uint32_t *>(lds);
    if (!lay.lu_global)
        for (uint32_t i = tid; i < B.words; i += nt) tl[i] = blob[i];
    double *block = lds + table_lds_words(B.words, lay) / 2;
    const WaveTemplate W = wave_template(B, lay.lu_global ? blob : tl);
    WaveLane s = wave_lane(W, lay, block);
    if (lay.lu_global) wave_bind_lu_scratch(s, W, lu_g + (size_t)l * lane_scratch_words(W.cap, W.n_tape));
    double *prm_lds = block + lay.prm;
    for (uint32_t i = tid; i < W.n_prm_lds; i += nt) prm_lds[i] = prm_lanes[(size_t)l * W.n_prm + i];
    s.gprm = prm_lanes + (size_t)l * W.n_prm;
    for (uint32_t i = tid; i < n; i += nt) s.x[i] = x_glob[(size_t)i * N + l];
    __syncthreads();
    const ReduceCache rc = wave_reduce_cache(W, tid);
    for (uint32_t e = tid; e < W.n_inst; e += nt) wave_eval_inst<true, VA>(W, s, e, false, 0.0);
    __syncthreads();
    wave_reduce(W, s, rc, tid, nt);
    for (uint32_t e = tid; e < n + nnz; e += nt) wave_reduce_charge(W, s, e);
    __syncthreads();
    if (tid == 0) s.g[W.diag[0]] += 1.0;
    __syncthreads();
    double *const Fr = block + lay.words;
    double *const Fi = Fr + (size_t)n * n;
    const double br = tid < n ? drive[tid] : 0.0, bi = tid < n ? drive[n + tid] : 0.0;
    for (uint32_t k = 0; k < n_freq; k++) {
        const double w = omega[k];
        for (uint32_t i = tid; i < n * n; i += nt) {
            Fr[i] = 0.0;
            Fi[i] = 0.0;
        }
        __syncthreads();
        if (tid < n)
            for (uint32_t p = W.P.col_ptr[tid]; p < W.P.col_ptr[tid + 1]; p++) {
                Fr[tid * n + W.P.row_idx[p]] = s.g[p];
                Fi[tid * n + W.P.row_idx[p]] = w * s.cc[p];
            }
        uint32_t perm_k = tid;
        bool singular = false;
        __syncthreads();
        for (uint32_t c = 0; c < n; c++) {
            double best = -1.0;
            if (tid >= c && tid < n) {
                best = __builtin_fabs(Fr[c * n + tid]) + __builtin_fabs(Fi[c * n + tid]);
                if (!(best == best)) best = __builtin_huge_val();
            }
            uint32_t at = tid;
            for (int sh = 32; sh >= 1; sh >>= 1) {
                const double ob = __shfl_xor(best, sh, 64);
                const uint32_t oa = __shfl_xor(at, sh, 64);
                if (ob > best || (ob == best && oa < at)) {
                    best = ob;
                    at = oa;
                }
            }
            if (!(best > 0.0) || !__builtin_isfinite(best)) {
                singular = true;
                break;
            }
            if (at != c) {
                if (tid < n) {
                    const double ar = Fr[tid * n + c], ai = Fi[tid * n + c];
                    Fr[tid * n + c] = Fr[tid * n + at];
                    Fi[tid * n + c] = Fi[tid * n + at];
                    Fr[tid * n + at] = ar;
                    Fi[tid * n + at] = ai;
                }
                const uint32_t pc = __shfl(perm_k, c, 64), pa = __shfl(perm_k, at, 64);
                if (tid == c) perm_k = pa;
                if (tid == at) perm_k = pc;
            }
            __syncthreads();
            // multipliers L(r, c) = A(r, c) / A(c, c)
            const double pr = Fr[c * n + c], pi = Fi[c * n + c];
            const double den = pr * pr + pi * pi;
            if (tid > c && tid < n) {
                const double ar = Fr[c * n + tid], ai = Fi[c * n + tid];
                Fr[c * n + tid] = (ar * pr + ai * pi) / den;
                Fi[c * n + tid] = (ai * pr - ar * pi) / den;
            }
            __syncthreads();
            if (tid > c && tid < n) {
                const double ur = Fr[tid * n + c], ui = Fi[tid * n + c];
                if (ur != 0.0 || ui != 0.0)
                    for (uint32_t r = c + 1; r < n; r++) {
                        const double lr = Fr[c * n + r], li = Fi[c * n + r];
                        Fr[tid * n + r] -= lr * ur - li * ui;
                        Fi[tid * n + r] -= lr * ui + li * ur;
                    }
            }
            __syncthreads();
        }
        double yr = __shfl(br, (int)perm_k, 64), yi = __shfl(bi, (int)perm_k, 64);
        if (!singular) {
            for (uint32_t c = 0; c + 1 < n; c++) {
                const double cr = __shfl(yr, (int)c, 64), ci = __shfl(yi, (int)c, 64);
                if (tid > c && tid < n) {
                    const double lr = Fr[c * n + tid], li = Fi[c * n + tid];
                    yr -= lr * cr - li * ci;
                    yi -= lr * ci + li * cr;
                }
            }
            for (uint32_t c = n; c-- > 0;) {
                if (tid == c) {
                    const double pr = Fr[c * n + c], pi = Fi[c * n + c];
                    const double den = pr * pr + pi * pi;
                    const double zr = (yr * pr + yi * pi) / den, zi = (yi * pr - yr * pi) / den;
                    yr = zr;
                    yi = zi;
                }
                const double cr = __shfl(yr, (int)c, 64), ci = __shfl(yi, (int)c, 64);
                if (tid < c) {
                    const double ur = Fr[c * n + tid], ui = Fi[c * n + tid];
                    yr -= ur * cr - ui * ci;
                    yi -= ur * ci + ui * cr;
                }
            }
        } else {
            yr = __builtin_nan("");
            yi = yr;
        }
        if (tid < n) {
            out[((size_t)(2 * k) * n + tid) * N + l] = yr;
            out[((size_t)(2 * k + 1) * n + tid) * N + l] = yi;
        }
        __syncthreads();
    }
}

// ------------------------------------------- the allocation-free replay

// Scratch of espice_lanes_tab_refactor_solve_device: the tape's index words
// followed by every lane's lx, ux, udiag, w and y planes. One array per GPU,
// held until the process ends; host code reaches it only through the
// address hipGetSymbolAddress gives per device (scratch_base).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / 8];

__global__ __launch_bounds__(256) void fill_zero_kernel(double *p, size_t count)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// ------------------------------------------------------------------ batch

struct espice_lanes {
    const Template *t = nullptr;
    uint32_t N = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_w0 = nullptr, ev_w1 = nullptr;
    std::vector<double> uprm[K_NKIND];  // host, lane-major (hand-written kinds)
    // compiled kinds: per-lane values of the parameters set_param moved,
    // keyed (kind, instance, parameter); the rest are the template's
    std::map<uint64_t, std::vector<double>> va_over;
    bool prepared = false;
    size_t lds_max = 64 * 1024;  // dynamic LDS a workgroup may take
    bool has_va = false;         // the template holds compiled kinds
    int va_mask = 0;             // which (va_bit), the kernels' instantiation
    // device pattern
    DevBuf<uint32_t> col_ptr, row_idx, q, tri_pos, diag, nodes[K_NKIND], slots[K_NKIND];
    DevBuf<uint8_t> cur_row;
    // device lane planes
    DevBuf<double> prm[K_NKIND], lim[K_NKIND], x, x_old, dx, g, rhs;
    DevBuf<uint32_t> status, iters;
    HostBuf<uint32_t> h_status, h_iters;
    HostBuf<double> h_x;
    uint32_t *d_status = nullptr, *d_iters = nullptr;  // their device addresses
    double *d_x = nullptr;
    // lane solver
    DevBuf<uint8_t> st, void_col, scaled_pivot;
    DevBuf<uint32_t> cnt, pinv, lp, up, li, ui, prow, void_slots, flag, topo, stack, pstack, want, want_of, prev;
    DevBuf<double> lx, ux, udiag, w, y, rscale, vcopy, tri, probe;
    uint32_t cap = 0;
    // workgroup-per-lane tables
    WaveHost wh;
    DevBuf<uint32_t> w_blob;
    DevBuf<double> prm_lanes;
    DevBuf<uint32_t> pred;      // the template's predicted first factor (PredTape)
    DevBuf<uint32_t> pred_hit;  // per lane, the last Newton launch's first factor: 0 none, 1 predicted, 2 serial
    DevBuf<double> lu_g;        // per lane, the serial L/U arrays of a lu_global layout (HBM)
    bool has_pred = false;
    WaveLayout lay{};
    size_t lds_bytes = 0;
    bool use_wave = false;
    std::vector<double> prm_host;
    // the operating point's homotopy launches: per-lane controls, the gmin
    // stamps, and the prepared-constant words holding a compiled kind's
    // $simparam("gmin") (found once per prepare)
    DevBuf<LaneHom> hom_lane;
    DevBuf<uint32_t> hom_stamp, hom_gslot;
    DevBuf<double> hom_xprev;
    uint32_t n_gslot = 0;
    bool hom_ready = false;

    ~espice_lanes()
    {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ev_w0) (void)hipEventDestroy(ev_w0);
        if (ev_w1) (void)hipEventDestroy(ev_w1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    Batch batch() const
    {
        Batch B;
        B.N = N;
        B.P = Pattern{t->n, t->nnz, col_ptr.p, row_idx.p, q.p, t->tridiagonal ? tri_pos.p : nullptr, t->diag_fill, cap};
        B.diag = diag.p;
        B.cur_row = cur_row.p;
        B.n_order = t->n_order;
        for (int k = 0; k < K_NKIND; k++) {
            B.order[k] = t->order[k];
            B.count[k] = t->count[k];
            B.kslots[k] = t->kslots[k];
            B.nodes[k] = nodes[k].p;
            B.slots[k] = slots[k].p;
            B.prm[k] = prm[k].p;
            B.lim[k] = lim[k].p;
        }
        B.x = x.p;
        B.x_old = x_old.p;
        B.dx = dx.p;
        B.g = g.p;
        B.rhs = rhs.p;
        B.status = status.p;
        B.iters = iters.p;
        B.lu = LaneLUBuf{st.p,    cnt.p,    pinv.p, lp.p,    up.p,      li.p,         ui.p,   prow.p,
                         void_slots.p, flag.p, topo.p, stack.p, pstack.p, want.p, want_of.p, prev.p,
                         void_col.p, scaled_pivot.p, lx.p, ux.p, udiag.p, w.p, y.p, rscale.p,
                         vcopy.p, tri.p, probe.p};
        return B;
    }
};

// The workgroup-per-lane kernels of each set of compiled kinds (va_bit):
// their launches, and their addresses for hipFuncSetAttribute.
template <int V>
static void launch_analyze(espice_lanes *L, const Newton &o)
{
    hipLaunchKernelGGL(analyze_wave_kernel<V>, dim3(1), dim3(64), L->lds_bytes, L->stream, L->wh.at, L->w_blob.p,
                       L->lay, L->prm_lanes.p, o, L->pred.p, L->lu_g.p);
}

template <int V>
static void launch_newton_dense(espice_lanes *L, const Newton &o, bool want_x, const Homotopy &hom)
{
    hipLaunchKernelGGL((newton_wave_kernel<true, V>), dim3(L->N), dim3(64), L->lds_bytes, L->stream, L->wh.at,
                       L->w_blob.p, L->lay, L->prm_lanes.p, L->x.p, L->N, o, L->status.p, L->iters.p,
                       L->has_pred ? L->pred.p : nullptr, L->pred_hit.p, L->d_status, L->d_iters,
                       want_x ? L->d_x : nullptr, L->lu_g.p, hom, MixedCtl{});
}

template <int V>
static void launch_newton_mixed(espice_lanes *L, const Newton &o, size_t lds, const MixedCtl &mix)
{
    hipLaunchKernelGGL((newton_wave_kernel<true, V, true>), dim3(L->N), dim3(64), lds, L->stream, L->wh.at,
                       L->w_blob.p, L->lay, L->prm_lanes.p, L->x.p, L->N, o, L->status.p, L->iters.p,
                       L->has_pred ? L->pred.p : nullptr, L->pred_hit.p, L->d_status, L->d_iters, nullptr,
                       L->lu_g.p, Homotopy{}, mix);
}

template <int V, bool MIXED = false>
static void launch_tran(espice_lanes *L, uint32_t N, const WaveLayout &lay, size_t lds, const Newton &op,
                        const Newton &step, const TranOpts &to, const MeasPass *passes, uint32_t n_passes,
                        TranOut *outs, MeasOut *meas, double *rec_t, double *rec_x, const MixedCtl &mix)
{
    hipLaunchKernelGGL((tran_wave_kernel<V, MIXED>), dim3(N), dim3(64), lds, L->stream, L->wh.at, L->w_blob.p, lay,
                       L->prm_lanes.p, L->x.p, N, op, step, to, passes, n_passes, outs, meas, rec_t, rec_x,
                       L->has_pred && lay.dense != ESP_NONE ? L->pred.p : nullptr, L->lu_g.p, mix);
}

static void (*const LAUNCH_ANALYZE[VA_SETS])(espice_lanes *, const Newton &) = {
    launch_analyze<0>, launch_analyze<1>, launch_analyze<2>, launch_analyze<3>};
static void (*const LAUNCH_NEWTON[VA_SETS])(espice_lanes *, const Newton &, bool, const Homotopy &) = {
    launch_newton_dense<0>, launch_newton_dense<1>, launch_newton_dense<2>, launch_newton_dense<3>};
// the mixed-precision Newton is built for the hand-written kinds and for
// templates holding PSP103 (the op-amp); the other sets are refused
static void (*const LAUNCH_NEWTON_MIXED[VA_SETS])(espice_lanes *, const Newton &, size_t, const MixedCtl &) = {
    launch_newton_mixed<0>, launch_newton_mixed<1>, nullptr, nullptr};
static const void *const NEWTON_MIXED_KERNEL[VA_SETS] = {
    (const void *)newton_wave_kernel<true, 0, true>, (const void *)newton_wave_kernel<true, 1, true>, nullptr,
    nullptr};
template <int V>
static void launch_ac(espice_lanes *L, const WaveLayout &lay, size_t lds, const double *omega, uint32_t n_freq,
                      const double *drive, double *out)
{
    hipLaunchKernelGGL(ac_wave_kernel<V>, dim3(L->N), dim3(64), lds, L->stream, L->wh.at, L->w_blob.p, lay,
                       L->prm_lanes.p, L->x.p, L->N, omega, n_freq, drive, out, L->lu_g.p);
}
static void (*const AC_LAUNCH[VA_SETS])(espice_lanes *, const WaveLayout &, size_t, const double *, uint32_t,
                                        const double *, double *) = {launch_ac<0>, launch_ac<1>, launch_ac<2>,
                                                                     launch_ac<3>};
static const void *const AC_KERNEL[VA_SETS] = {(const void *)ac_wave_kernel<0>, (const void *)ac_wave_kernel<1>,
                                               (const void *)ac_wave_kernel<2>, (const void *)ac_wave_kernel<3>};

typedef void (*TranLaunch)(espice_lanes *, uint32_t, const WaveLayout &, size_t, const Newton &, const Newton &,
                           const TranOpts &, const MeasPass *, uint32_t, TranOut *, MeasOut *, double *, double *,
                           const MixedCtl &);
static const TranLaunch TRAN_LAUNCH[VA_SETS] = {launch_tran<0>, launch_tran<1>, launch_tran<2>, launch_tran<3>};
// the mixed-precision transient, for the hand-written kinds and templates
// holding PSP103 (the op-amp)
static const TranLaunch TRAN_LAUNCH_MIXED[VA_SETS] = {launch_tran<0, true>, launch_tran<1, true>, nullptr,
                                                      nullptr};
static const void *const TRAN_MIXED_KERNEL[VA_SETS] = {(const void *)tran_wave_kernel<0, true>,
                                                       (const void *)tran_wave_kernel<1, true>, nullptr, nullptr};
static const void *const ANALYZE_KERNEL[VA_SETS] = {
    (const void *)analyze_wave_kernel<0>, (const void *)analyze_wave_kernel<1>, (const void *)analyze_wave_kernel<2>,
    (const void *)analyze_wave_kernel<3>};
static const void *const NEWTON_KERNEL[VA_SETS] = {
    (const void *)newton_wave_kernel<true, 0>, (const void *)newton_wave_kernel<true, 1>,
    (const void *)newton_wave_kernel<true, 2>, (const void *)newton_wave_kernel<true, 3>};
static const void *const TRAN_KERNEL[VA_SETS] = {
    (const void *)tran_wave_kernel<0>, (const void *)tran_wave_kernel<1>, (const void *)tran_wave_kernel<2>,
    (const void *)tran_wave_kernel<3>};

namespace {

template <class T>
hipError_t upload(DevBuf<T> &d, const std::vector<T> &h)
{
    hipError_t e = d.alloc(h.size());
    if (e != hipSuccess) return e;
    if (h.empty()) return hipSuccess;
    return hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
}

template <class T>
hipError_t zeroed(DevBuf<T> &d, size_t count)
{
    hipError_t e = d.alloc(count);
    if (e != hipSuccess) return e;
    return hipMemset(d.p, 0, d.n * sizeof(T));
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

espice_status espice_lanes_error_message(char *buffer, size_t capacity, size_t *required)
{
    if (required) *required = g_error.size() + 1;
    if (capacity < g_error.size() + 1) return ESPICE_BUFFER_TOO_SMALL;
    if (!buffer) return ESPICE_INVALID_ARGUMENT;
    std::memcpy(buffer, g_error.c_str(), g_error.size() + 1);
    return ESPICE_OK;
}

espice_status espice_lanes_template_create(espice_lanes_template **out)
{
    if (!out) return fail(ESPICE_INVALID_ARGUMENT, "NullOutput");
    *out = new (std::nothrow) espice_lanes_template();
    return *out ? ESPICE_OK : fail(ESPICE_OUT_OF_MEMORY, "OutOfMemory");
}

espice_status espice_lanes_template_destroy(espice_lanes_template *t)
{
    delete t;
    return ESPICE_OK;
}

espice_status espice_lanes_template_node(espice_lanes_template *t, const char *name, uint32_t *row)
{
    if (!t || !name || !row) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (t->t.frozen) return fail(ESPICE_INVALID_ARGUMENT, "TemplateFrozen");
    *row = t->t.node(name);
    return ESPICE_OK;
}

espice_status espice_lanes_template_add(espice_lanes_template *t, uint32_t kind, const uint32_t *ports,
                                        const double *params, size_t param_count, uint32_t *instance)
{
    if (!t || !ports || !instance || kind >= K_NKIND) return fail(ESPICE_INVALID_ARGUMENT, "InvalidDevice");
    if (param_count != (size_t)KIND_UK[kind] || (param_count && !params)) return fail(ESPICE_INVALID_ARGUMENT, "ParamCount");
    if (t->t.add((int)kind, ports, params, instance) != 0) return fail(ESPICE_INVALID_ARGUMENT, "InvalidDevice");
    return ESPICE_OK;
}

espice_status espice_lanes_param_count(uint32_t kind, uint32_t *count)
{
    if (!count || kind >= K_NKIND) return fail(ESPICE_INVALID_ARGUMENT, "InvalidDevice");
    *count = (uint32_t)KIND_UK[kind];
    return ESPICE_OK;
}

espice_status espice_lanes_param_index(uint32_t kind, const char *name, uint32_t *index)
{
    if (!name || !index || kind >= K_NKIND) return fail(ESPICE_INVALID_ARGUMENT, "InvalidDevice");
    if (!kind_is_va((int)kind)) return fail(ESPICE_INVALID_ARGUMENT, "UnknownParam");
    const char *const *names = va_param_names((int)kind);
    for (int k = 0; k < KIND_UK[kind]; k++)
        if (strcasecmp(names[k], name) == 0) {
            *index = (uint32_t)k;
            return ESPICE_OK;
        }
    for (int k = 0; k < va_nalias((int)kind); k++)
        if (strcasecmp(va_alias_names((int)kind)[k], name) == 0) {
            *index = (uint32_t)va_alias_target((int)kind)[k];
            return ESPICE_OK;
        }
    return fail(ESPICE_INVALID_ARGUMENT, "UnknownParam");
}

espice_status espice_lanes_template_freeze(espice_lanes_template *t)
{
    if (!t) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (t->t.freeze() != 0) return fail(ESPICE_FAILED, "OutOfWorkspace");
    return ESPICE_OK;
}

espice_status espice_lanes_template_gmin_stamps(const espice_lanes_template *t, uint32_t *col, uint32_t *slot)
{
    if (!t || !col || !slot) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (!t->t.frozen) return fail(ESPICE_INVALID_ARGUMENT, "NotFrozen");
    std::vector<uint32_t> st;
    t->t.gmin_stamps(st);
    const uint32_t n = t->t.n;
    std::memcpy(col, st.data(), 4 * (size_t)n);
    std::memcpy(slot, st.data() + n, 4 * (size_t)n);
    return ESPICE_OK;
}

espice_status espice_lanes_template_info(const espice_lanes_template *t, uint32_t *n, uint32_t *nnz, uint32_t *tridiagonal)
{
    if (!t || !t->t.frozen) return fail(ESPICE_INVALID_ARGUMENT, "NotFrozen");
    if (n) *n = t->t.n;
    if (nnz) *nnz = t->t.nnz;
    if (tridiagonal) *tridiagonal = t->t.tridiagonal ? 1 : 0;
    return ESPICE_OK;
}

espice_status espice_lanes_template_pattern(const espice_lanes_template *t, uint32_t *col_ptr, uint32_t *row_idx,
                                            uint32_t *diag, uint8_t *current_row, uint32_t *q)
{
    if (!t || !t->t.frozen) return fail(ESPICE_INVALID_ARGUMENT, "NotFrozen");
    const Template &T = t->t;
    if (col_ptr) std::memcpy(col_ptr, T.col_ptr.data(), 4 * (T.n + 1));
    if (row_idx) std::memcpy(row_idx, T.row_idx.data(), 4 * (size_t)T.nnz);
    if (diag) std::memcpy(diag, T.diag.data(), 4 * (size_t)T.n);
    if (current_row) std::memcpy(current_row, T.cur_row.data(), T.n);
    if (q) std::memcpy(q, T.q.data(), 4 * (size_t)T.n);
    return ESPICE_OK;
}

espice_status espice_lanes_create(const espice_lanes_template *tt, uint32_t n_lanes, espice_lanes **out)
{
    if (!tt || !out || n_lanes == 0) return fail(ESPICE_INVALID_ARGUMENT, "InvalidLanes");
    if (!tt->t.frozen) return fail(ESPICE_INVALID_ARGUMENT, "NotFrozen");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) return fail(ESPICE_FAILED, "NoGpu");
    std::unique_ptr<espice_lanes> L(new (std::nothrow) espice_lanes());
    if (!L) return fail(ESPICE_OUT_OF_MEMORY, "OutOfMemory");
    const Template &T = tt->t;
    L->t = &T;
    L->N = n_lanes;
    const size_t N = n_lanes, n = T.n, nnz = T.nnz;
    HIP_TRY(hipStreamCreate(&L->stream));
    HIP_TRY(hipEventCreate(&L->ev0));
    HIP_TRY(hipEventCreate(&L->ev1));
    HIP_TRY(hipEventCreate(&L->ev_w0));
    HIP_TRY(hipEventCreate(&L->ev_w1));
    for (int k = 0; k < K_NKIND; k++) {
        if (kind_is_va(k)) continue;
        L->uprm[k].resize((size_t)T.count[k] * KIND_UK[k] * N);
        for (size_t i = 0; i < T.count[k]; i++)
            for (int p = 0; p < KIND_UK[k]; p++)
                std::fill_n(&L->uprm[k][(i * KIND_UK[k] + p) * N], N, T.uprm[k][i * KIND_UK[k] + p]);
    }
    HIP_TRY(upload(L->col_ptr, T.col_ptr));
    HIP_TRY(upload(L->row_idx, T.row_idx));
    HIP_TRY(upload(L->q, T.q));
    HIP_TRY(upload(L->tri_pos, T.tri_pos));
    HIP_TRY(upload(L->diag, T.diag));
    HIP_TRY(upload(L->cur_row, T.cur_row));
    for (int k = 0; k < K_NKIND; k++) {
        HIP_TRY(upload(L->nodes[k], T.nodes[k]));
        HIP_TRY(upload(L->slots[k], T.slots[k]));
        // the thread-per-lane kernel's per-kind planes; compiled kinds run on
        // the workgroup-per-lane kernels only
        const size_t pc = kind_is_va(k) ? 0 : (size_t)T.count[k];
        HIP_TRY(zeroed(L->prm[k], pc * KIND_PK[k] * N));
        HIP_TRY(zeroed(L->lim[k], pc * KIND_NU[k] * N));
    }
    HIP_TRY(zeroed(L->x, n * N));
    HIP_TRY(zeroed(L->pred_hit, N));
    HIP_TRY(L->h_status.alloc(N));
    HIP_TRY(L->h_iters.alloc(N));
    HIP_TRY(L->h_x.alloc(n * N));
    HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void **>(&L->d_status), L->h_status.p, 0));
    HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void **>(&L->d_iters), L->h_iters.p, 0));
    HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void **>(&L->d_x), L->h_x.p, 0));
    HIP_TRY(zeroed(L->x_old, n * N));
    HIP_TRY(zeroed(L->dx, n * N));
    HIP_TRY(zeroed(L->g, (nnz + 1) * N));
    HIP_TRY(zeroed(L->rhs, (n + 1) * N));
    HIP_TRY(zeroed(L->status, N));
    HIP_TRY(zeroed(L->iters, N));
    // L + U of any factor with n pivots fit n (n - 1) entries.
    L->cap = (uint32_t)std::max<size_t>(n * (n - 1), 1);
    const size_t cap = L->cap;
    HIP_TRY(zeroed(L->st, 3 * N));
    HIP_TRY(zeroed(L->cnt, 4 * N));
    HIP_TRY(zeroed(L->pinv, n * N));
    HIP_TRY(zeroed(L->lp, (n + 1) * N));
    HIP_TRY(zeroed(L->up, (n + 1) * N));
    HIP_TRY(zeroed(L->li, cap * N));
    HIP_TRY(zeroed(L->ui, cap * N));
    HIP_TRY(zeroed(L->prow, nnz * N));
    HIP_TRY(zeroed(L->void_slots, nnz * N));
    for (DevBuf<uint32_t> *b : {&L->flag, &L->topo, &L->stack, &L->pstack, &L->want, &L->want_of, &L->prev})
        HIP_TRY(zeroed(*b, n * N));
    HIP_TRY(zeroed(L->void_col, n * N));
    HIP_TRY(zeroed(L->scaled_pivot, n * N));
    HIP_TRY(zeroed(L->lx, cap * N));
    HIP_TRY(zeroed(L->ux, cap * N));
    for (DevBuf<double> *b : {&L->udiag, &L->w, &L->y, &L->rscale}) HIP_TRY(zeroed(*b, n * N));
    HIP_TRY(zeroed(L->vcopy, nnz * N));
    HIP_TRY(zeroed(L->tri, 3 * n * N));
    HIP_TRY(zeroed(L->probe, 3 * n * N));
    build_wave(T, cap, L->wh);
    HIP_TRY(upload(L->w_blob, L->wh.blob));
    HIP_TRY(zeroed(L->prm_lanes, (size_t)L->wh.n_prm * N));
    {
        // sizes only: the lane block's layout does not depend on where the
        // template tables live
        const WaveTemplate W = wave_template(L->wh.at, L->wh.blob.data());
        // templates with compiled kinds keep the serial factor's L/U arrays
        // in HBM: their lane blocks then fit more workgroups per CU
        bool va = false;
        for (int k = 0; k < K_NKIND; k++) va = va || (kind_is_va(k) && T.count[k] > 0);
        L->lay.lu_global = va ? 1u : 0u;
        L->lds_bytes = 4 * table_lds_words(L->wh.at.words, L->lay) + 8 * wave_words(W, L->lay);
        HIP_TRY(L->lu_g.alloc(va ? lane_scratch_words(cap, W.n_tape) * N : 1));
        int dev = 0, lds = 0;
        HIP_TRY(hipGetDevice(&dev));
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
            L->lds_max = std::min<size_t>((size_t)lds, 160 * 1024);
        L->use_wave = L->lds_bytes <= L->lds_max;
    }
    for (int k = 0; k < K_NKIND; k++) L->has_va = L->has_va || (kind_is_va(k) && T.count[k] > 0);
    for (int k = 0; k < K_NKIND; k++) L->va_mask |= T.count[k] > 0 ? va_bit(k) : 0;
    if (L->use_wave && L->lds_bytes > 64 * 1024) {
        // a lane block past the default 64 KiB of dynamic LDS (gfx950 gives
        // a workgroup up to 160 KiB): the kernels this template launches
        // are allowed what it needs
        const int m = (int)L->lds_bytes;
        std::string bad;
        const void *ks[] = {NEWTON_KERNEL[L->va_mask], ANALYZE_KERNEL[L->va_mask]};
        for (int i = 0; i < 2; i++)
            if (hipFuncSetAttribute(ks[i], hipFuncAttributeMaxDynamicSharedMemorySize, m) != hipSuccess) {
                (void)hipGetLastError();
                bad += bad.empty() ? "" : ",";
                bad += std::to_string(i);
            }
        if (!bad.empty()) return fail(ESPICE_FAILED, ("LdsAttribute " + bad + " at " + std::to_string(m)).c_str());
    }
    if (L->has_va && (!L->use_wave || L->lay.dense == ESP_NONE)) return fail(ESPICE_FAILED, "LaneTooLarge");
    const uint32_t blocks = (n_lanes + 63) / 64;
    hipLaunchKernelGGL(init_lane_state_kernel, dim3(blocks), dim3(64), 0, L->stream, L->st.p, n_lanes,
                       (uint8_t)(T.tridiagonal ? 1 : 0));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(L->stream));
    *out = L.release();
    return ESPICE_OK;
}

espice_status espice_lanes_destroy(espice_lanes *l)
{
    delete l;
    return ESPICE_OK;
}

espice_status espice_lanes_reset(espice_lanes *L)
{
    if (!L) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    // the workgroup-per-lane kernels build every lane's solver state in LDS
    // at launch; only the thread-per-lane solver keeps it across launches
    if (L->use_wave) return ESPICE_OK;
    const uint32_t blocks = (L->N + 63) / 64;
    hipLaunchKernelGGL(init_lane_state_kernel, dim3(blocks), dim3(64), 0, L->stream, L->st.p, L->N,
                       (uint8_t)(L->t->tridiagonal ? 1 : 0));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(L->cnt.p, 0, L->cnt.n * sizeof(uint32_t), L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    return ESPICE_OK;
}

espice_status espice_lanes_set_param(espice_lanes *L, uint32_t kind, uint32_t instance, uint32_t param, const double *values)
{
    if (!L || !values || kind >= K_NKIND) return fail(ESPICE_INVALID_ARGUMENT, "InvalidDevice");
    if (instance >= L->t->count[kind] || param >= (uint32_t)KIND_UK[kind]) return fail(ESPICE_INVALID_ARGUMENT, "InvalidParam");
    if (kind_is_va((int)kind)) {
        L->va_over[((uint64_t)kind << 56) | ((uint64_t)instance << 24) | param].assign(values, values + L->N);
        L->prepared = false;
        return ESPICE_OK;
    }
    std::memcpy(&L->uprm[kind][((size_t)instance * KIND_UK[kind] + param) * L->N], values, sizeof(double) * L->N);
    L->prepared = false;
    return ESPICE_OK;
}

__global__ __launch_bounds__(256) void set_strided_kernel(double *p, size_t stride, size_t offset, double v, uint32_t N)
{
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < N) p[(size_t)l * stride + offset] = v;
}

espice_status espice_lanes_set_source(espice_lanes *L, uint32_t kind, uint32_t instance, double value)
{
    if (!L) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if ((kind != K_VSRC && kind != K_ISRC) || instance >= L->t->count[kind])
        return fail(ESPICE_INVALID_ARGUMENT, "InvalidDevice");
    if (!L->prepared) return fail(ESPICE_INVALID_ARGUMENT, "NotPrepared");
    if (!(value == value)) return fail(ESPICE_INVALID_ARGUMENT, "InvalidParam");
    const size_t N = L->N, uk = KIND_UK[kind], pk = KIND_PK[kind];
    // the user value (a later prepare keeps it) and its prepared copy, which
    // is the value itself once given (models/vsource.va:216-221)
    for (size_t l = 0; l < N; l++) L->uprm[kind][((size_t)instance * uk + VSRC_DC) * N + l] = value;
    const InstEntry *ie = reinterpret_cast<const InstEntry *>(L->wh.inst_bytes.data());
    const dim3 grid((uint32_t)((N + 255) / 256));
    for (uint32_t e = 0; e < L->wh.n_inst; e++) {
        if (ie[e].kind != kind || ie[e].id != instance) continue;
        const size_t word = ie[e].prm + SRCP_DC;
        for (size_t l = 0; l < N; l++) L->prm_host[l * L->wh.n_prm + word] = value;
        hipLaunchKernelGGL(set_strided_kernel, grid, dim3(256), 0, L->stream, L->prm_lanes.p, (size_t)L->wh.n_prm,
                           word, value, (uint32_t)N);
    }
    if (L->prm[kind].p)
        hipLaunchKernelGGL(set_strided_kernel, grid, dim3(256), 0, L->stream, L->prm[kind].p, (size_t)1,
                           ((size_t)instance * pk + SRCP_DC) * N, value, (uint32_t)N);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(L->stream));
    return ESPICE_OK;
}

// A compiled kind's setup on every lane and instance, on the host's
// threads, straight into the lane-contiguous constants. Every lane must
// keep the template's collapse (its internal unknowns' rows), as a
// diode's RS must keep its sign.
static espice_status prepare_va(espice_lanes *L, int kind)
{
    const Template &T = *L->t;
    const size_t N = L->N, n_prm = L->wh.n_prm;
    const int uk = KIND_UK[kind], pk = KIND_PK[kind];
    const InstEntry *ie = reinterpret_cast<const InstEntry *>(L->wh.inst_bytes.data());
    std::vector<uint32_t> prm_of(T.count[kind], 0);
    for (uint32_t e = 0; e < L->wh.n_inst; e++)
        if (ie[e].kind == kind) prm_of[ie[e].id] = ie[e].prm;
    // the template's collapse flags per instance, from its own parameters
    const int nc = va_ncoll(kind), cb = va_coll_base(kind);
    std::vector<double> flags((size_t)T.count[kind] * (nc ? nc : 1));
    {
        std::vector<double> p(pk);
        for (uint32_t i = 0; i < T.count[kind]; i++) {
            if (va_setup(kind, &T.uprm[kind][(size_t)i * uk], p.data()) != 0)
                return fail(ESPICE_INVALID_ARGUMENT, "ParamRange");
            for (int c = 0; c < nc; c++) flags[(size_t)i * nc + c] = p[cb + c];
        }
    }
    std::vector<std::vector<std::pair<int, const double *>>> over(T.count[kind]);
    for (const auto &kv : L->va_over) {
        if ((int)(kv.first >> 56) != kind) continue;
        over[(kv.first >> 24) & 0xffffffffu].push_back({(int)(kv.first & 0xffffffu), kv.second.data()});
    }
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = (unsigned)std::min<size_t>(hw, std::max<size_t>(1, N / 256));
    std::vector<int> bad(nt, 0);
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; w++)
        pool.emplace_back([&, w]() {
            std::vector<double> u(uk), p(pk);
            for (size_t l = w; l < N; l += nt)
                for (uint32_t i = 0; i < T.count[kind]; i++) {
                    std::memcpy(u.data(), &T.uprm[kind][(size_t)i * uk], 8 * (size_t)uk);
                    for (const auto &o : over[i]) u[o.first] = o.second[l];
                    if (va_setup(kind, u.data(), p.data()) != 0) {
                        bad[w] = 1;
                        continue;
                    }
                    for (int c = 0; c < nc; c++)
                        if (p[cb + c] != flags[(size_t)i * nc + c]) bad[w] = 2;
                    std::memcpy(&L->prm_host[l * n_prm + prm_of[i]], p.data(), 8 * (size_t)pk);
                }
        });
    for (auto &t : pool) t.join();
    for (int b : bad) {
        if (b == 1) return fail(ESPICE_INVALID_ARGUMENT, "ParamRange");
        if (b == 2) return fail(ESPICE_INVALID_ARGUMENT, "CollapseTopology");
    }
    return ESPICE_OK;
}

espice_status espice_lanes_prepare(espice_lanes *L)
{
    if (!L) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    const size_t N = L->N;
    L->prm_host.assign((size_t)L->wh.n_prm * N, 0.0);
    for (int kind = 0; kind < K_NKIND; kind++) {
        const uint32_t cnt = L->t->count[kind];
        const int uk = KIND_UK[kind], pk = KIND_PK[kind];
        if (kind_is_va(kind)) {
            if (espice_status e = prepare_va(L, kind)) return e;
            continue;
        }
        std::vector<double> host((size_t)cnt * pk * N);
        double u[PARAM_MAX], p[PARAM_MAX];
        for (uint32_t i = 0; i < cnt; i++)
            for (size_t l = 0; l < N; l++) {
                for (int k = 0; k < uk; k++) u[k] = L->uprm[kind][((size_t)i * uk + k) * N + l];
                // a diode's ai row is the template's topology: RS may move
                // per lane but not across zero (diode.va:364-365)
                // SFFM on a current source follows isrcload.c's own slots and
                // a dc fill from the .tran defaults (isource.va, builder.zig
                // dcFromWaveform); not restated here
                if (kind == K_ISRC && u[VSRC_WAVE] == 5.0) return fail(ESPICE_INVALID_ARGUMENT, "IsourceSffmUnsupported");
                if (kind == K_DIODE) {
                    const uint32_t *nd = &L->t->nodes[K_DIODE][(size_t)i * 3];
                    if ((u[DIO_RS] > 0.0) != (nd[2] != nd[0])) return fail(ESPICE_INVALID_ARGUMENT, "DiodeRsTopology");
                }
                prep_kind(kind, u, p);
                for (int k = 0; k < pk; k++) host[((size_t)i * pk + k) * N + l] = p[k];
            }
        if (!host.empty())
            HIP_TRY(hipMemcpyAsync(L->prm[kind].p, host.data(), host.size() * 8, hipMemcpyHostToDevice, L->stream));
        HIP_TRY(hipStreamSynchronize(L->stream));
        // the same constants per lane contiguous, in serial instance order
        const InstEntry *ie = reinterpret_cast<const InstEntry *>(L->wh.inst_bytes.data());
        for (uint32_t e = 0; e < L->wh.n_inst; e++) {
            if (ie[e].kind != kind) continue;
            for (size_t l = 0; l < N; l++)
                for (int k = 0; k < pk; k++)
                    L->prm_host[l * L->wh.n_prm + ie[e].prm + k] = host[((size_t)ie[e].id * pk + k) * N + l];
        }
    }
    if (!L->prm_host.empty())
        HIP_TRY(hipMemcpy(L->prm_lanes.p, L->prm_host.data(), L->prm_host.size() * 8, hipMemcpyHostToDevice));
    // the symbolic analysis every lane's first factor tries (wave_lu.h
    // dense_first_factor), from lane 0's cold start
    L->has_pred = false;
    if (L->use_wave && L->lay.dense != ESP_NONE && !L->t->tridiagonal) {
        const PredTape T = pred_tape(L->t->n, L->t->nnz);
        if (L->pred.n < T.words) HIP_TRY(L->pred.alloc(T.words));
        const Newton o{1e-3, 1e-12, 1e-6, 1e-9, 0.0, 100, 1};
        LAUNCH_ANALYZE[L->va_mask](L, o);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(L->stream));
        L->has_pred = true;
    }
    L->prepared = true;
    L->hom_ready = false;
    return ESPICE_OK;
}

espice_status espice_lanes_warm_start(espice_lanes *L, espice_mlp *m, const double *params, double *kernel_ms)
{
    if (!L || !m || !params) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    HIP_TRY(hipEventRecord(L->ev0, L->stream));
    if (espice_status e = mlp_warm_start(m, params, L->N, L->x.p, L->t->n, L->N, L->stream)) return e;
    HIP_TRY(hipEventRecord(L->ev1, L->stream));
    HIP_TRY(hipEventSynchronize(L->ev1));
    if (kernel_ms) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, L->ev0, L->ev1));
        *kernel_ms = ms;
    }
    return ESPICE_OK;
}

espice_status espice_lanes_upload_x(espice_lanes *L, const double *x)
{
    if (!L || !x) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    HIP_TRY(hipMemcpy(L->x.p, x, sizeof(double) * L->t->n * L->N, hipMemcpyHostToDevice));
    return ESPICE_OK;
}

espice_status espice_lanes_download_x(espice_lanes *L, double *x)
{
    if (!L || !x) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    HIP_TRY(hipMemcpy(x, L->x.p, sizeof(double) * L->t->n * L->N, hipMemcpyDeviceToHost));
    return ESPICE_OK;
}

// The static Newton kernel of every lane on the lanes' stream, bracketed by
// ev0/ev1; with want_x the workgroup-per-lane kernel also writes x to the
// mapped host copy.
static espice_status launch_newton(espice_lanes *L, const espice_newton_opts *opts, bool want_x,
                                   const Homotopy &hom = Homotopy{})
{
    if (!L->prepared) return fail(ESPICE_INVALID_ARGUMENT, "NotPrepared");
    if (opts->max_iter == 0) return fail(ESPICE_INVALID_ARGUMENT, "MaxIter");
    if (hom.lane != nullptr && !L->use_wave) return fail(ESPICE_INVALID_ARGUMENT, "HomotopyLaneTooLarge");
    Newton o{opts->reltol, opts->abstol, opts->vntol, opts->residual_tol, opts->gmin, opts->max_iter, opts->init_fix};
    HIP_TRY(hipEventRecord(L->ev0, L->stream));
    if (L->use_wave) {
        if (L->lay.dense != ESP_NONE)
            LAUNCH_NEWTON[L->va_mask](L, o, want_x, hom);
        else
            hipLaunchKernelGGL((newton_wave_kernel<false, 0>), dim3(L->N), dim3(64), L->lds_bytes, L->stream, L->wh.at,
                               L->w_blob.p, L->lay, L->prm_lanes.p, L->x.p, L->N, o, L->status.p, L->iters.p,
                               nullptr, nullptr, L->d_status, L->d_iters, want_x ? L->d_x : nullptr, L->lu_g.p, hom,
                               MixedCtl{});
    } else {
        const Batch B = L->batch();
        const uint32_t blocks = (L->N + 63) / 64;
        hipLaunchKernelGGL(newton_op_kernel, dim3(blocks), dim3(64), 0, L->stream, B, o);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(L->ev1, L->stream));
    return ESPICE_OK;
}

// Status, iterations and (when x is given) the solution back through the
// pinned staging, then the one synchronisation of the solve.
static espice_status finish_newton(espice_lanes *L, espice_lane_result *out, double *x, double *newton_ms)
{
    const size_t N = L->N, nx = (size_t)L->t->n * N;
    if (!L->use_wave) {
        HIP_TRY(hipMemcpyAsync(L->h_status.p, L->status.p, 4 * N, hipMemcpyDeviceToHost, L->stream));
        HIP_TRY(hipMemcpyAsync(L->h_iters.p, L->iters.p, 4 * N, hipMemcpyDeviceToHost, L->stream));
        if (x) HIP_TRY(hipMemcpyAsync(L->h_x.p, L->x.p, 8 * nx, hipMemcpyDeviceToHost, L->stream));
    }
    HIP_TRY(hipStreamSynchronize(L->stream));
    if (newton_ms) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, L->ev0, L->ev1));
        *newton_ms = ms;
    }
    for (size_t l = 0; l < N; l++) out[l] = espice_lane_result{L->h_status.p[l], L->h_iters.p[l]};
    if (x) std::memcpy(x, L->h_x.p, 8 * nx);
    return ESPICE_OK;
}

espice_status espice_lanes_shape(const espice_lanes *L, uint32_t *n, uint32_t *n_lanes)
{
    if (!L) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (n) *n = L->t->n;
    if (n_lanes) *n_lanes = L->N;
    return ESPICE_OK;
}

// The homotopy tables of a prepared batch: the template's gmin stamps and,
// per compiled instance, the prepared constants that carry the model's
// gmin. Those are found on lane 0 by a second setup at another gmin: a
// constant that moves must equal the gmin in both setups (a copy of
// $simparam("gmin"), as psp103's and bsim4va's are), so writing the
// homotopy's gmin there is the reference's setup rerun.
static espice_status homotopy_tables(espice_lanes *L)
{
    if (L->hom_ready) return ESPICE_OK;
    const Template &T = *L->t;
    std::vector<uint32_t> st;
    T.gmin_stamps(st);
    HIP_TRY(upload(L->hom_stamp, st));
    std::vector<uint32_t> gslot;
    const InstEntry *ie = reinterpret_cast<const InstEntry *>(L->wh.inst_bytes.data());
    constexpr double PROBE = 0x1p-15;
    for (uint32_t e = 0; e < L->wh.n_inst; e++) {
        const int kind = ie[e].kind;
        if (!kind_is_va(kind)) continue;
        const int uk = KIND_UK[kind], pk = KIND_PK[kind];
        std::vector<double> u(&T.uprm[kind][(size_t)ie[e].id * uk], &T.uprm[kind][(size_t)ie[e].id * uk] + uk);
        for (const auto &kv : L->va_over)
            if ((int)(kv.first >> 56) == kind && ((kv.first >> 24) & 0xffffffffu) == ie[e].id)
                u[kv.first & 0xffffffu] = kv.second[0];
        std::vector<double> p0(pk), p1(pk);
        if (va_setup(kind, u.data(), p0.data()) != 0 || va_setup(kind, u.data(), p1.data(), PROBE) != 0)
            return fail(ESPICE_INVALID_ARGUMENT, "ParamRange");
        for (int k = 0; k < pk; k++) {
            if (std::memcmp(&p0[k], &p1[k], 8) == 0) continue;
            if (p0[k] != VA_SIMPARAM_GMIN || p1[k] != PROBE) return fail(ESPICE_INVALID_ARGUMENT, "HomotopyGminDerived");
            gslot.push_back(ie[e].prm + (uint32_t)k);
        }
    }
    L->n_gslot = (uint32_t)gslot.size();
    if (gslot.empty()) gslot.push_back(0);  // a valid buffer; n_gslot says none
    HIP_TRY(upload(L->hom_gslot, gslot));
    HIP_TRY(L->hom_lane.alloc(L->N));
    L->hom_ready = true;
    return ESPICE_OK;
}

espice_status espice_lanes_newton_homotopy(espice_lanes *L, const espice_newton_opts *opts,
                                           const espice_lane_homotopy *lanes, const double *x_prev,
                                           espice_lane_result *out, double *kernel_ms)
{
    static_assert(sizeof(espice_lane_homotopy) == sizeof(LaneHom), "the C-ABI control is LaneHom");
    if (!L || !opts || !lanes || !out) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (!L->prepared) return fail(ESPICE_INVALID_ARGUMENT, "NotPrepared");
    bool predict = false;
    for (size_t l = 0; l < L->N; l++) predict = predict || (lanes[l].flags & ESPICE_HOM_PREDICT);
    if (predict && !x_prev) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (espice_status e = homotopy_tables(L)) return e;
    HIP_TRY(hipMemcpyAsync(L->hom_lane.p, lanes, sizeof(LaneHom) * L->N, hipMemcpyHostToDevice, L->stream));
    if (predict) {
        HIP_TRY(L->hom_xprev.alloc((size_t)L->t->n * L->N));
        HIP_TRY(hipMemcpyAsync(L->hom_xprev.p, x_prev, sizeof(double) * L->t->n * L->N, hipMemcpyHostToDevice,
                               L->stream));
    }
    const Homotopy hom{L->hom_lane.p, L->hom_stamp.p, L->hom_gslot.p, L->n_gslot,
                       predict ? L->hom_xprev.p : nullptr};
    if (espice_status e = launch_newton(L, opts, false, hom)) return e;
    if (espice_status e = finish_newton(L, out, nullptr, kernel_ms)) return e;
    for (size_t l = 0; l < L->N; l++)
        if (!(lanes[l].flags & ESPICE_HOM_ACTIVE)) out[l] = espice_lane_result{0, 0};
    return ESPICE_OK;
}

espice_status espice_lanes_newton_mixed(espice_lanes *L, const espice_newton_opts *opts, uint32_t max_refine,
                                        espice_lane_result *out, uint32_t *refine, double *kernel_ms)
{
    if (!L || !opts || !out) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (!L->prepared) return fail(ESPICE_INVALID_ARGUMENT, "NotPrepared");
    if (opts->max_iter == 0 || max_refine == 0) return fail(ESPICE_INVALID_ARGUMENT, "MaxIter");
    if (!L->use_wave || L->lay.dense == ESP_NONE) return fail(ESPICE_FAILED, "LaneTooLarge");
    if (!LAUNCH_NEWTON_MIXED[L->va_mask]) return fail(ESPICE_FAILED, "MixedUnsupported");
    const size_t lds = L->lds_bytes + 8 * mixed_words(L->t->n, L->t->nnz);
    if (lds > L->lds_max) return fail(ESPICE_FAILED, "LaneTooLarge");
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(NEWTON_MIXED_KERNEL[L->va_mask], hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    DevBuf<uint32_t> d_mix;
    HIP_TRY(d_mix.alloc(2 * (size_t)L->N));
    Newton o{opts->reltol, opts->abstol, opts->vntol, opts->residual_tol, opts->gmin, opts->max_iter, opts->init_fix};
    HIP_TRY(hipEventRecord(L->ev0, L->stream));
    LAUNCH_NEWTON_MIXED[L->va_mask](L, o, lds, MixedCtl{max_refine, d_mix.p});
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(L->ev1, L->stream));
    if (refine)
        HIP_TRY(hipMemcpyAsync(refine, d_mix.p, 8 * (size_t)L->N, hipMemcpyDeviceToHost, L->stream));
    return finish_newton(L, out, nullptr, kernel_ms);
}

espice_status espice_lanes_newton_op(espice_lanes *L, const espice_newton_opts *opts, espice_lane_result *out, double *kernel_ms)
{
    if (!L || !opts || !out) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (espice_status e = launch_newton(L, opts, false)) return e;
    return finish_newton(L, out, nullptr, kernel_ms);
}

espice_status espice_lanes_newton_learned(espice_lanes *L, espice_mlp *m, const double *params,
                                          const espice_newton_opts *opts, espice_lane_result *out, double *x,
                                          double *warm_ms, double *newton_ms)
{
    if (!L || !m || !params || !opts || !out) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (!L->prepared) return fail(ESPICE_INVALID_ARGUMENT, "NotPrepared");
    // every check launch_newton makes comes before the inference is queued:
    // the inference reads the network's mapped input buffer in place, so no
    // error may return while it could still run
    if (opts->max_iter == 0) return fail(ESPICE_INVALID_ARGUMENT, "MaxIter");
    hipEvent_t w0 = L->ev_w0, w1 = L->ev_w1;
    HIP_TRY(hipEventRecord(w0, L->stream));
    if (espice_status e = mlp_warm_start(m, params, L->N, L->x.p, L->t->n, L->N, L->stream)) {
        (void)hipStreamSynchronize(L->stream);
        return e;
    }
    HIP_TRY(hipEventRecord(w1, L->stream));
    if (espice_status e = launch_newton(L, opts, x != nullptr)) {
        (void)hipStreamSynchronize(L->stream);
        return e;
    }
    if (espice_status e = finish_newton(L, out, x, newton_ms)) return e;
    if (warm_ms) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, w0, w1));
        *warm_ms = ms;
    }
    return ESPICE_OK;
}

espice_status espice_lanes_first_factors(espice_lanes *L, uint32_t *predicted, uint32_t *serial)
{
    if (!L || !predicted || !serial) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    std::vector<uint32_t> h(L->N);
    HIP_TRY(hipMemcpy(h.data(), L->pred_hit.p, 4 * (size_t)L->N, hipMemcpyDeviceToHost));
    *predicted = 0;
    *serial = 0;
    for (uint32_t v : h) {
        *predicted += v == 1 ? 1 : 0;
        *serial += v == 2 ? 1 : 0;
    }
    return ESPICE_OK;
}

static_assert(KIND_UK[K_VSRC] == ESPICE_LANES_VSOURCE_PARAMS && KIND_UK[K_ISRC] == ESPICE_LANES_ISOURCE_PARAMS &&
                  KIND_UK[K_RES] == ESPICE_LANES_RESISTOR_PARAMS && KIND_UK[K_CAP] == ESPICE_LANES_CAPACITOR_PARAMS &&
                  KIND_UK[K_DIODE] == ESPICE_LANES_DIODE_PARAMS && KIND_UK[K_MOS1] == ESPICE_LANES_MOS1_PARAMS,
              "user parameter counts match include/espice_lanes.h");
static_assert(DIO_RS == ESPICE_LANES_DIODE_RS, "diode RS slot matches include/espice_lanes.h");
static_assert(sizeof(espice_tran_opts) == sizeof(TranOpts), "espice_tran_opts mirrors TranOpts");
static_assert(sizeof(espice_meas_pass) == sizeof(MeasPass), "espice_meas_pass mirrors MeasPass");
static_assert(sizeof(espice_meas_result) == sizeof(MeasOut), "espice_meas_result mirrors MeasOut");
static_assert(sizeof(espice_tran_result) == sizeof(TranOut), "espice_tran_result mirrors TranOut");

// Whether every MOSFET of every lane has zero charge: mos1.va charges only
// with an oxide (TOX given and > 0, :164), overlaps CGSO/CGDO/CGBO, or
// junction capacitance (CBD/CBS, or CJ and CJSW with nonzero AD/AS/PD/PS,
// :253-255).
static bool tran_charges_zero(const espice_lanes &L)
{
    auto zero = [](double v) { return v != v || v == 0.0; };
    auto none = [](double v) { return v != v || v <= 0.0; };
    const size_t N = L.N;
    for (uint32_t i = 0; i < L.t->count[K_MOS1]; i++)
        for (size_t l = 0; l < N; l++) {
            const double *u = &L.uprm[K_MOS1][(size_t)i * KIND_UK[K_MOS1] * N + l];
            auto at = [&](int k) { return u[(size_t)k * N]; };
            if (!none(at(6))) return false;                                      // TOX
            if (!zero(at(22)) || !zero(at(23)) || !zero(at(24))) return false;  // CGSO CGDO CGBO
            if (!none(at(14)) || !none(at(15))) return false;                   // CBD CBS
            if (!zero(at(16)) && !(zero(at(31)) && zero(at(32)))) return false;  // CJ with AD/AS
            if (!zero(at(17)) && !(zero(at(33)) && zero(at(34)))) return false;  // CJSW with PD/PS
        }
    return true;
}

espice_status espice_lanes_tran(espice_lanes *L, const espice_tran_opts *opts, const espice_meas_pass *passes,
                                uint32_t n_passes, espice_tran_result *out, espice_meas_result *meas, double *times,
                                double *rows, double *kernel_ms)
{
    if (!L || !opts || !out || (n_passes && (!passes || !meas))) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (!L->prepared) return fail(ESPICE_INVALID_ARGUMENT, "NotPrepared");
    const uint32_t n = L->t->n, N = L->N;
    const size_t cap = opts->record_capacity;
    if (cap && (!times || !rows)) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (!(opts->t_stop > 0) || !(opts->dt_init > 0) || !(opts->dt_max > 0) || opts->method > 2)
        return fail(ESPICE_INVALID_ARGUMENT, "InvalidTran");
    for (uint32_t k = 0; k < n_passes; k++)
        if (passes[k].kind > 4 || passes[k].row >= n) return fail(ESPICE_INVALID_ARGUMENT, "InvalidMeasure");
    // a Find reads the Cross before it as its WHEN; the device measures one
    // as it goes only when that WHEN has no TD=, FROM= or LAST count (the
    // cards the reference measures as it goes, docs/using/cli.md:186-197)
    for (uint32_t k = 0; k < n_passes; k++)
        if (passes[k].kind == 4) {
            const espice_meas_pass &w = passes[k - (k > 0 ? 1 : 0)];
            if (k == 0 || w.kind != 0 || w.td != 0 || w.from != 0 || w.rise == -2 || w.fall == -2 || w.cross == -2)
                return fail(ESPICE_INVALID_ARGUMENT, "InvalidMeasure");
        }
    // The Meyer and junction charges of mos1.va are not staged yet: a lane
    // whose MOSFETs would carry charge is refused
    // rather than run as static; with every such parameter zero (or not
    // given) their charges are identically zero and the lane is exact.
    if (!tran_charges_zero(*L)) return fail(ESPICE_FAILED, "TranChargesUnsupported");
    WaveLayout lay{};
    lay.lu_global = L->lay.lu_global;
    const WaveTemplate W = wave_template(L->wh.at, L->wh.blob.data());
    const size_t words = wave_words(W, lay, true, tran_extra_words(W.n_inst, n_passes));
    const bool mixed = opts->mixed_refine != 0;
    if (mixed && !TRAN_LAUNCH_MIXED[L->va_mask]) return fail(ESPICE_FAILED, "MixedUnsupported");
    const size_t lds = 4 * table_lds_words(L->wh.at.words, lay) + 8 * words + (mixed ? 8 * mixed_words(n, L->t->nnz) : 0);
    if (lay.dense == ESP_NONE || lds > L->lds_max) return fail(ESPICE_FAILED, "LaneTooLarge");
    DevBuf<MeasPass> d_pass;
    DevBuf<TranOut> d_out;
    DevBuf<MeasOut> d_meas;
    DevBuf<double> d_t, d_x;
    HIP_TRY(d_pass.alloc(n_passes));
    HIP_TRY(d_out.alloc(N));
    HIP_TRY(d_meas.alloc((size_t)N * n_passes));
    HIP_TRY(d_t.alloc(cap ? (size_t)N * cap : 1));
    HIP_TRY(d_x.alloc(cap ? (size_t)N * cap * n : 1));
    if (n_passes)
        HIP_TRY(hipMemcpyAsync(d_pass.p, passes, sizeof(MeasPass) * n_passes, hipMemcpyHostToDevice, L->stream));
    TranOpts to;
    std::memcpy(&to, opts, sizeof to);
    Newton op{};
    op.reltol = opts->op_reltol > 0.0 ? opts->op_reltol : opts->reltol;
    op.abstol = opts->abstol;
    op.vntol = opts->vntol;
    op.residual_tol = 1e-9;
    op.gmin = 0.0;
    op.max_iter = opts->max_iter;
    op.init_fix = opts->op_warm ? 0 : 1;
    Newton step = op;
    step.residual_tol = inf_();
    step.init_fix = 0;
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(mixed ? TRAN_MIXED_KERNEL[L->va_mask] : TRAN_KERNEL[L->va_mask],
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    HIP_TRY(hipEventRecord(L->ev0, L->stream));
    (mixed ? TRAN_LAUNCH_MIXED : TRAN_LAUNCH)[L->va_mask](L, N, lay, lds, op, step, to, d_pass.p, n_passes, d_out.p,
                                                         d_meas.p, d_t.p, d_x.p, MixedCtl{opts->mixed_refine, nullptr});
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(L->ev1, L->stream));
    HIP_TRY(hipMemcpyAsync(out, d_out.p, sizeof(TranOut) * N, hipMemcpyDeviceToHost, L->stream));
    if (n_passes)
        HIP_TRY(hipMemcpyAsync(meas, d_meas.p, sizeof(MeasOut) * N * n_passes, hipMemcpyDeviceToHost, L->stream));
    if (cap) {
        HIP_TRY(hipMemcpyAsync(times, d_t.p, sizeof(double) * N * cap, hipMemcpyDeviceToHost, L->stream));
        HIP_TRY(hipMemcpyAsync(rows, d_x.p, sizeof(double) * N * cap * n, hipMemcpyDeviceToHost, L->stream));
    }
    HIP_TRY(hipStreamSynchronize(L->stream));
    if (kernel_ms) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, L->ev0, L->ev1));
        *kernel_ms = ms;
    }
    return ESPICE_OK;
}

espice_status espice_lanes_template_ac_drive(const espice_lanes_template *t, uint32_t kind, uint32_t instance,
                                            double mag, double phase_deg, double *drive)
{
    if (!t || !drive) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (!t->t.frozen) return fail(ESPICE_INVALID_ARGUMENT, "NotFrozen");
    if ((kind != K_VSRC && kind != K_ISRC) || instance >= t->t.count[kind])
        return fail(ESPICE_INVALID_ARGUMENT, "InvalidDevice");
    const uint32_t n = t->t.n;
    const double rad = phase_deg * (3.14159265358979323846 / 180.0);
    const double re = mag * std::cos(rad), im = mag * std::sin(rad);
    // (pos, neg, re, im) as the reference's excitation (builder.zig:802-817,
    // :1146, :1158): a V card's branch row is its neg
    const uint32_t *nd = &t->t.nodes[kind][(size_t)instance * KIND_NU[kind]];
    const uint32_t pos = kind == K_VSRC ? 0 : nd[0], neg = kind == K_VSRC ? nd[2] : nd[1];
    if (pos != 0) {
        drive[pos] -= re;
        drive[n + pos] -= im;
    }
    if (neg != 0) {
        drive[neg] += re;
        drive[n + neg] += im;
    }
    return ESPICE_OK;
}

espice_status espice_lanes_ac(espice_lanes *L, const double *omega, uint32_t n_freq, const double *drive,
                              double *out, double *kernel_ms)
{
    if (!L || !omega || !drive || !out) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (!L->prepared) return fail(ESPICE_INVALID_ARGUMENT, "NotPrepared");
    if (n_freq == 0) return ESPICE_OK;
    if (!tran_charges_zero(*L)) return fail(ESPICE_FAILED, "AcChargesUnsupported");
    const uint32_t n = L->t->n, N = L->N;
    WaveLayout lay{};
    lay.lu_global = L->lay.lu_global;
    const WaveTemplate W = wave_template(L->wh.at, L->wh.blob.data());
    const size_t words = wave_words(W, lay, true, tran_extra_words(W.n_inst, 0));
    const size_t lds = 4 * table_lds_words(L->wh.at.words, lay) + 8 * words + 16 * (size_t)n * n;
    if (lay.dense == ESP_NONE || lds > L->lds_max) return fail(ESPICE_FAILED, "LaneTooLarge");
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(AC_KERNEL[L->va_mask], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    DevBuf<double> d_w, d_drive, d_out;
    HIP_TRY(d_w.alloc(n_freq));
    HIP_TRY(d_drive.alloc(2 * (size_t)n));
    HIP_TRY(d_out.alloc((size_t)n_freq * 2 * n * N));
    HIP_TRY(hipMemcpyAsync(d_w.p, omega, 8 * (size_t)n_freq, hipMemcpyHostToDevice, L->stream));
    HIP_TRY(hipMemcpyAsync(d_drive.p, drive, 16 * (size_t)n, hipMemcpyHostToDevice, L->stream));
    HIP_TRY(hipEventRecord(L->ev0, L->stream));
    AC_LAUNCH[L->va_mask](L, lay, lds, d_w.p, n_freq, d_drive.p, d_out.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(L->ev1, L->stream));
    HIP_TRY(hipMemcpyAsync(out, d_out.p, 8 * (size_t)n_freq * 2 * n * N, hipMemcpyDeviceToHost, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    if (kernel_ms) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, L->ev0, L->ev1));
        *kernel_ms = ms;
    }
    return ESPICE_OK;
}

espice_status espice_lanes_tab_build(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const double *vals,
                                     espice_lane_tab *tab, uint32_t *idx, size_t idx_capacity, size_t *len)
{
    if (!col_ptr || !row_idx || !vals || !tab || !len || n == 0) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    const uint32_t nnz = col_ptr[n];
    std::vector<uint32_t> q(n);
    if (column_order(n, col_ptr, row_idx, q.data()) != 0) return fail(ESPICE_FAILED, "OutOfWorkspace");
    if (is_tridiag(n, col_ptr, row_idx)) {
        // ownTape: the Thomas engine keeps no LU, so there is no tape unless
        // its factor fails and demotes (direct.zig:359-362).
        std::vector<uint32_t> pos(3 * (size_t)n, ESP_NONE);
        for (uint32_t j = 0; j < n; j++)
            for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
                uint32_t r = row_idx[p];
                if (r == j) pos[n + j] = p;
                else if (r == j + 1) pos[j + 1] = p;
                else if (j > 0 && r + 1 == j) pos[2 * n + r] = p;
            }
        std::vector<double> tri(3 * (size_t)n);
        Pattern P{n, nnz, col_ptr, row_idx, q.data(), pos.data(), 0, 0};
        LaneLU s{};
        s.bp = LV<double>{tri.data(), 1};
        s.mul = s.bp.at(n);
        s.cv = s.bp.at(2 * (size_t)n);
        if (tri_factor(P, s, vals) == 0) return fail(ESPICE_FAILED, "NoLaneTape");
    }
    const uint64_t df = std::max<uint64_t>(2 * (chol_count(n, col_ptr, row_idx, q.data(), UINT64_MAX) - n), nnz);
    const size_t cap = std::max<size_t>((size_t)n * (n - 1), 1);
    std::vector<uint8_t> st(3, 0), vcol(n), spiv(n);
    std::vector<uint32_t> cnt(4, 0), pinv(n), lp(n + 1), up(n + 1), li(cap), ui(cap), prow(nnz + 1), vslots(nnz + 1);
    std::vector<uint32_t> flag(n), topo(n), stack(n), pstack(n), want(n), want_of(n), prev(n);
    std::vector<double> lx(cap), ux(cap), udiag(n), w(n, 0.0), y(n), rscale(n), vcopy(nnz + 1), tri(3 * (size_t)n),
        probe(3 * (size_t)n);
    LaneLUBuf buf{st.data(),    cnt.data(),    pinv.data(), lp.data(),   up.data(),      li.data(),      ui.data(),
                  prow.data(),  vslots.data(), flag.data(), topo.data(), stack.data(),   pstack.data(),  want.data(),
                  want_of.data(), prev.data(), vcol.data(), spiv.data(), lx.data(),      ux.data(),      udiag.data(),
                  w.data(),     y.data(),      rscale.data(), vcopy.data(), tri.data(), probe.data()};
    Pattern P{n, nnz, col_ptr, row_idx, q.data(), nullptr, df, (uint32_t)cap};
    LaneLU s = lane_lu(buf, 1, 0, n);
    if (direct_factor(P, s, vals) != 0) return fail(ESPICE_FAILED, "SingularMatrix");
    const uint32_t nl = cnt[C_NL], nu = cnt[C_NU], nv = cnt[C_NVOID];
    espice_lane_tab T{};
    T.n = n;
    T.nnz = nnz;
    T.nnz_l = nl;
    T.nnz_u = nu;
    T.growth = GROWTH_LIMIT;
    T.wide = 1 / PIVOT_TOL;
    T.n_void_slots = nv;
    const size_t total = (size_t)(n + 1) + nnz + n + n + (n + 1) + nl + (n + 1) + nu + nv + n;
    *len = total;
    if (idx_capacity < total || !idx) return fail(ESPICE_BUFFER_TOO_SMALL, "BufferTooSmall");
    uint32_t at = 0;
    auto part = [&](uint32_t &field, const uint32_t *src, size_t cntw) {
        field = at;
        if (cntw) std::memcpy(idx + at, src, 4 * cntw);
        at += (uint32_t)cntw;
    };
    part(T.col_ptr, col_ptr, n + 1);
    part(T.prow, prow.data(), nnz);
    part(T.q, q.data(), n);
    part(T.pinv, pinv.data(), n);
    part(T.lp, lp.data(), n + 1);
    part(T.li, li.data(), nl);
    part(T.up, up.data(), n + 1);
    part(T.ui, ui.data(), nu);
    part(T.void_slots, vslots.data(), nv);
    T.flags = at;
    for (uint32_t k = 0; k < n; k++) idx[at + k] = (vcol[k] ? 1u : 0u) | (spiv[k] ? 2u : 0u);
    *tab = T;
    return ESPICE_OK;
}

espice_status espice_lanes_tab_refactor_solve(const espice_lane_tab *tab, const uint32_t *idx, size_t idx_len,
                                              uint32_t n_lanes, const double *a, const double *rhs, double *dx,
                                              uint8_t *fail_out, double *kernel_ms)
{
    if (!tab || !idx || !a || !rhs || !dx || !fail_out || n_lanes == 0) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    const espice_lane_tab &T = *tab;
    // Every word the bodies index must lie inside idx: checked here, once,
    // before anything runs on the device.
    const size_t need = (size_t)T.flags + T.n;
    if (need > idx_len || T.col_ptr + T.n + 1 > idx_len || T.prow + T.nnz > idx_len || T.li + T.nnz_l > idx_len ||
        T.ui + T.nnz_u > idx_len || T.void_slots + T.n_void_slots > idx_len)
        return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t k = 0; k <= T.n; k++)
        if (idx[T.lp + k] > T.nnz_l || idx[T.up + k] > T.nnz_u || idx[T.col_ptr + k] > T.nnz)
            return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t k = 0; k < T.n; k++)
        if (idx[T.q + k] >= T.n || idx[T.pinv + k] >= T.n) return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t p = 0; p < T.nnz; p++)
        if (idx[T.prow + p] >= T.n) return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t p = 0; p < T.nnz_l; p++)
        if (idx[T.li + p] >= T.n) return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t p = 0; p < T.nnz_u; p++)
        if (idx[T.ui + p] >= T.n) return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t v = 0; v < T.n_void_slots; v++)
        if (idx[T.void_slots + v] >= T.nnz) return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) return fail(ESPICE_FAILED, "NoGpu");
    const size_t N = n_lanes, n = T.n;
    DevBuf<uint32_t> d_idx;
    DevBuf<double> d_a, d_rhs, d_dx, d_lx, d_ux, d_udiag, d_w, d_y;
    DevBuf<uint8_t> d_fail;
    HIP_TRY(d_idx.alloc(idx_len));
    HIP_TRY(hipMemcpy(d_idx.p, idx, 4 * idx_len, hipMemcpyHostToDevice));
    HIP_TRY(d_a.alloc((size_t)T.nnz * N));
    HIP_TRY(hipMemcpy(d_a.p, a, 8 * (size_t)T.nnz * N, hipMemcpyHostToDevice));
    HIP_TRY(d_rhs.alloc(n * N));
    HIP_TRY(hipMemcpy(d_rhs.p, rhs, 8 * n * N, hipMemcpyHostToDevice));
    HIP_TRY(zeroed(d_dx, n * N));
    HIP_TRY(zeroed(d_lx, std::max<size_t>(T.nnz_l, 1) * N));
    HIP_TRY(zeroed(d_ux, std::max<size_t>(T.nnz_u, 1) * N));
    HIP_TRY(zeroed(d_udiag, n * N));
    HIP_TRY(zeroed(d_w, n * N));
    HIP_TRY(zeroed(d_y, n * N));
    HIP_TRY(zeroed(d_fail, N));
    LaneTab t{T.n, T.nnz, T.nnz_l, T.nnz_u, T.growth, T.wide, T.col_ptr, T.prow, T.q, T.pinv, T.lp, T.li, T.up, T.ui,
              T.flags, T.void_slots, T.n_void_slots, 0};
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(tab_refactor_kernel, dim3((n_lanes + 63) / 64), dim3(64), 0, 0, t, d_idx.p, n_lanes, d_a.p, d_rhs.p,
                       d_dx.p, d_fail.p, d_lx.p, d_ux.p, d_udiag.p, d_w.p, d_y.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, 0));
    HIP_TRY(hipDeviceSynchronize());
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel_ms) *kernel_ms = ms;
    HIP_TRY(hipMemcpy(dx, d_dx.p, 8 * n * N, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(fail_out, d_fail.p, N, hipMemcpyDeviceToHost));
    return ESPICE_OK;
}

// The scratch array's device address on the current GPU, looked up once per
// device and kept.
static espice_status scratch_base(double **out)
{
    static std::mutex mu;
    static double *base[64] = {};
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(ESPICE_FAILED, "DeviceIndex");
    std::lock_guard<std::mutex> guard(mu);
    if (!base[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        base[dev] = static_cast<double *>(p);
    }
    *out = base[dev];
    return ESPICE_OK;
}

espice_status espice_lanes_tab_refactor_solve_device(const espice_lane_tab *tab, const uint32_t *idx, size_t idx_len,
                                                     uint32_t n_lanes, const double *a, const double *rhs, double *dx,
                                                     uint8_t *fail_out, void *stream)
{
    if (!tab || !idx || !a || !rhs || !dx || !fail_out || n_lanes == 0) return fail(ESPICE_INVALID_ARGUMENT, "NullArgument");
    const espice_lane_tab &T = *tab;
    const size_t need = (size_t)T.flags + T.n;
    if (need > idx_len || T.col_ptr + T.n + 1 > idx_len || T.prow + T.nnz > idx_len || T.li + T.nnz_l > idx_len ||
        T.ui + T.nnz_u > idx_len || T.void_slots + T.n_void_slots > idx_len)
        return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t k = 0; k <= T.n; k++)
        if (idx[T.lp + k] > T.nnz_l || idx[T.up + k] > T.nnz_u || idx[T.col_ptr + k] > T.nnz)
            return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t k = 0; k < T.n; k++)
        if (idx[T.q + k] >= T.n || idx[T.pinv + k] >= T.n) return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t p = 0; p < T.nnz; p++)
        if (idx[T.prow + p] >= T.n) return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t p = 0; p < T.nnz_l; p++)
        if (idx[T.li + p] >= T.n) return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t p = 0; p < T.nnz_u; p++)
        if (idx[T.ui + p] >= T.n) return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    for (uint32_t v = 0; v < T.n_void_slots; v++)
        if (idx[T.void_slots + v] >= T.nnz) return fail(ESPICE_INVALID_ARGUMENT, "TabOutOfRange");
    const size_t N = n_lanes, n = T.n;
    // words of the scratch: the index words (two per double), then lx
    // (nnz_l), ux (nnz_u), udiag, w, y (n each) per lane
    const size_t idx_words = (idx_len + 1) / 2;
    const size_t nl = std::max<size_t>(T.nnz_l, 1), nu = std::max<size_t>(T.nnz_u, 1);
    const size_t words = idx_words + N * (nl + nu + 3 * n);
    if (words * 8 > NOALLOC_SCRATCH_BYTES) return fail(ESPICE_INVALID_ARGUMENT, "ScratchTooSmall");
    double *base = nullptr;
    if (espice_status e = scratch_base(&base)) return e;
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t *d_idx = reinterpret_cast<uint32_t *>(base);
    double *d_lx = base + idx_words, *d_ux = d_lx + N * nl, *d_udiag = d_ux + N * nu;
    double *d_w = d_udiag + N * n, *d_y = d_w + N * n;
    HIP_TRY(hipMemcpyAsync(d_idx, idx, 4 * idx_len, hipMemcpyHostToDevice, st));
    // w must start at zero (tab_refactor's scatter vector); a fill kernel,
    // not a memset, so the call can sit in a captured graph
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)std::min<size_t>((N * n + 255) / 256, 1024)), dim3(256), 0, st,
                       d_w, N * n);
    HIP_TRY(hipGetLastError());
    LaneTab t{T.n, T.nnz, T.nnz_l, T.nnz_u, T.growth, T.wide, T.col_ptr, T.prow, T.q, T.pinv, T.lp, T.li, T.up, T.ui,
              T.flags, T.void_slots, T.n_void_slots, 0};
    hipLaunchKernelGGL(tab_refactor_kernel, dim3((n_lanes + 63) / 64), dim3(64), 0, st, t, d_idx, n_lanes, a, rhs, dx,
                       fail_out, d_lx, d_ux, d_udiag, d_w, d_y);
    HIP_TRY(hipGetLastError());
    return ESPICE_OK;
}

}  // extern "C"
