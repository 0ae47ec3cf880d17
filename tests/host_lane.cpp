// Test-only harness: runs the lane engine's per-lane bodies (the exact
// ESP_HD code the gfx950 kernel runs, egspice_amd/csrc/engine.h) on the
// host, one lane at a time, so CPU tests can compare them with the oracle
// before any GPU run. This is the reference's own discipline of host-run
// device bodies against the oracles (src/solver/kernels/lane.zig:731-968).
// Never loaded by the engine.
#include <cstring>
#include <vector>

#include "../egspice_amd/csrc/engine.h"
#include "../egspice_amd/csrc/wave.h"
#include "../egspice_amd/csrc/wave_tran.h"

using namespace esp;

struct HL {
    Template *t;
    uint32_t N;
    std::vector<double> uprm[K_NKIND], prm[K_NKIND], lim[K_NKIND];
    std::vector<double> x, x_old, dx, g, rhs;
    std::vector<uint32_t> status, iters;
    std::vector<uint8_t> st, void_col, scaled_pivot, cur_row;
    std::vector<uint32_t> cnt, pinv, lp, up, li, ui, prow, void_slots, flag, topo, stack, pstack, want, want_of, prev;
    std::vector<double> lx, ux, udiag, w, y, rscale, vcopy, tri, probe;
    uint32_t cap;
};

extern "C" {

void *hl_template_new() { return new Template(); }
void hl_template_free(void *t) { delete (Template *)t; }
uint32_t hl_node(void *t, const char *name) { return ((Template *)t)->node(name); }
int hl_add(void *t, int kind, const uint32_t *ports, const double *params)
{
    uint32_t inst;
    return ((Template *)t)->add(kind, ports, params, &inst) == 0 ? (int)inst : -1;
}
int hl_freeze(void *t) { return ((Template *)t)->freeze(); }
uint32_t hl_n(void *t) { return ((Template *)t)->n; }
uint32_t hl_nnz(void *t) { return ((Template *)t)->nnz; }
void hl_pattern(void *tv, uint32_t *col_ptr, uint32_t *row_idx, uint32_t *diag, uint32_t *q)
{
    Template *t = (Template *)tv;
    std::memcpy(col_ptr, t->col_ptr.data(), 4 * (t->n + 1));
    std::memcpy(row_idx, t->row_idx.data(), 4 * (size_t)t->nnz);
    std::memcpy(diag, t->diag.data(), 4 * (size_t)t->n);
    std::memcpy(q, t->q.data(), 4 * (size_t)t->n);
}

void *hl_lanes_new(void *tv, uint32_t N)
{
    Template *t = (Template *)tv;
    HL *h = new HL();
    h->t = t;
    h->N = N;
    const size_t n = t->n, nnz = t->nnz;
    for (int k = 0; k < K_NKIND; k++) {
        h->uprm[k].resize((size_t)t->count[k] * KIND_UK[k] * N);
        for (size_t i = 0; i < t->count[k]; i++)
            for (int p = 0; p < KIND_UK[k]; p++)
                for (uint32_t l = 0; l < N; l++) h->uprm[k][(i * KIND_UK[k] + p) * N + l] = t->uprm[k][i * KIND_UK[k] + p];
        h->prm[k].assign((size_t)t->count[k] * KIND_PK[k] * N + 1, 0.0);
        h->lim[k].assign((size_t)t->count[k] * KIND_NU[k] * N + 1, 0.0);
    }
    h->x.assign(n * N, 0.0);
    h->x_old.assign(n * N, 0.0);
    h->dx.assign(n * N, 0.0);
    h->g.assign((nnz + 1) * N, 0.0);
    h->rhs.assign((n + 1) * N, 0.0);
    h->status.assign(N, 0);
    h->iters.assign(N, 0);
    h->cur_row = t->cur_row;
    h->cap = (uint32_t)std::max<size_t>(n * (n - 1), 1);
    const size_t cap = h->cap;
    h->st.assign(3 * N, 0);
    for (uint32_t l = 0; l < N; l++) h->st[(size_t)ST_TRI * N + l] = t->tridiagonal ? 1 : 0;
    h->cnt.assign(4 * N, 0);
    h->pinv.assign(n * N, 0);
    h->lp.assign((n + 1) * N, 0);
    h->up.assign((n + 1) * N, 0);
    h->li.assign(cap * N, 0);
    h->ui.assign(cap * N, 0);
    h->prow.assign(nnz * N + 1, 0);
    h->void_slots.assign(nnz * N + 1, 0);
    for (auto *v : {&h->flag, &h->topo, &h->stack, &h->pstack, &h->want, &h->want_of, &h->prev}) v->assign(n * N, 0);
    h->void_col.assign(n * N, 0);
    h->scaled_pivot.assign(n * N, 0);
    h->lx.assign(cap * N, 0.0);
    h->ux.assign(cap * N, 0.0);
    for (auto *v : {&h->udiag, &h->w, &h->y, &h->rscale}) v->assign(n * N, 0.0);
    h->vcopy.assign(nnz * N + 1, 0.0);
    h->tri.assign(3 * n * N, 0.0);
    h->probe.assign(3 * n * N, 0.0);
    return h;
}

void hl_lanes_free(void *h) { delete (HL *)h; }

void hl_set_param(void *hv, int kind, uint32_t inst, uint32_t k, const double *vals)
{
    HL *h = (HL *)hv;
    std::memcpy(&h->uprm[kind][((size_t)inst * KIND_UK[kind] + k) * h->N], vals, 8 * (size_t)h->N);
}

void hl_prepare(void *hv)
{
    HL *h = (HL *)hv;
    const size_t N = h->N;
    for (int kind = 0; kind < K_NKIND; kind++) {
        // the thread-per-lane bodies serve the hand-written kinds only
        if (kind_is_va(kind)) continue;
        double u[PARAM_MAX], p[PARAM_MAX];
        for (uint32_t i = 0; i < h->t->count[kind]; i++)
            for (size_t l = 0; l < N; l++) {
                for (int k = 0; k < KIND_UK[kind]; k++) u[k] = h->uprm[kind][((size_t)i * KIND_UK[kind] + k) * N + l];
                prep_kind(kind, u, p);
                for (int k = 0; k < KIND_PK[kind]; k++) h->prm[kind][((size_t)i * KIND_PK[kind] + k) * N + l] = p[k];
            }
    }
}

void hl_prm(void *hv, int kind, double *out)
{
    HL *h = (HL *)hv;
    std::memcpy(out, h->prm[kind].data(), 8 * ((size_t)h->t->count[kind] * KIND_PK[kind] * h->N));
}

void hl_newton_op(void *hv, const Newton *o, double *x, uint32_t *status, uint32_t *iters)
{
    HL *h = (HL *)hv;
    Template *t = h->t;
    const size_t N = h->N;
    std::memcpy(h->x.data(), x, 8 * t->n * N);
    Batch B;
    B.N = h->N;
    B.P = Pattern{t->n, t->nnz, t->col_ptr.data(), t->row_idx.data(), t->q.data(), t->tridiagonal ? t->tri_pos.data() : nullptr,
                  t->diag_fill, h->cap};
    B.diag = t->diag.data();
    B.cur_row = h->cur_row.data();
    B.n_order = t->n_order;
    for (int k = 0; k < K_NKIND; k++) {
        B.order[k] = t->order[k];
        B.count[k] = t->count[k];
        B.kslots[k] = t->kslots[k];
        B.nodes[k] = t->nodes[k].data();
        B.slots[k] = t->slots[k].data();
        B.prm[k] = h->prm[k].data();
        B.lim[k] = h->lim[k].data();
    }
    B.x = h->x.data();
    B.x_old = h->x_old.data();
    B.dx = h->dx.data();
    B.g = h->g.data();
    B.rhs = h->rhs.data();
    B.status = h->status.data();
    B.iters = h->iters.data();
    B.lu = LaneLUBuf{h->st.data(),    h->cnt.data(),   h->pinv.data(),     h->lp.data(),     h->up.data(),   h->li.data(),
                     h->ui.data(),    h->prow.data(),  h->void_slots.data(), h->flag.data(),  h->topo.data(), h->stack.data(),
                     h->pstack.data(), h->want.data(), h->want_of.data(),  h->prev.data(),   h->void_col.data(),
                     h->scaled_pivot.data(), h->lx.data(), h->ux.data(),   h->udiag.data(),  h->w.data(),    h->y.data(),
                     h->rscale.data(), h->vcopy.data(), h->tri.data(),     h->probe.data()};
    for (uint32_t l = 0; l < h->N; l++) lane_newton(B, l, *o);
    std::memcpy(x, h->x.data(), 8 * t->n * N);
    std::memcpy(status, h->status.data(), 4 * N);
    std::memcpy(iters, h->iters.data(), 4 * N);
}

// The workgroup-per-lane kernel's phase sequence (engine.hip
// newton_wave_kernel) replayed with the 64 threads of each phase run one
// after another; the barriers fall between phases.
void hl_newton_wave(void *hv, const Newton *o, double *x, uint32_t *status, uint32_t *iters)
{
    HL *h = (HL *)hv;
    Template *t = h->t;
    const uint32_t N = h->N, n = t->n, nnz = t->nnz, nt = 64;
    WaveHost wh;
    build_wave(*t, h->cap, wh);
    // the kernel's view: every table read from the packed blob
    const WaveTemplate W = wave_template(wh.at, wh.blob.data());
    WaveLayout lay{};
    size_t words = wave_words(W, lay);
    std::vector<double> block(words);
    const InstEntry *ie = W.inst;
    for (uint32_t l = 0; l < N; l++) {
        std::fill(block.begin(), block.end(), 0.0);
        WaveLane s = wave_lane(W, lay, block.data());
        double *prm = block.data() + lay.prm;
        for (uint32_t e = 0; e < wh.n_inst; e++) {
            int kind = ie[e].kind, pk = KIND_PK[kind];
            for (int k = 0; k < pk; k++) prm[ie[e].prm + k] = h->prm[kind][((size_t)ie[e].id * pk + k) * N + l];
        }
        for (uint32_t i = 0; i < n; i++) s.x[i] = s.x_old[i] = x[(size_t)i * N + l];
        wave_fresh_solver(W, s);
        bool lim_active = false;
        if (o->init_fix) {
            for (uint32_t e = 0; e < wh.n_inst; e++) wave_seed_inst(W, s, e);
            lim_active = true;
        }
        uint32_t bits = o->init_fix ? G_INITFIX : 0, st = 0, it = o->max_iter;
        double limit = 1e9, scaled = 0;
        for (uint32_t tick = 0; tick < o->max_iter; tick++) {
            for (uint32_t tid = 0; tid < nt; tid++)
                for (uint32_t e = tid; e < wh.n_inst; e += nt) wave_eval_inst(W, s, e, lim_active);
            for (uint32_t tid = 0; tid < nt; tid++) wave_reduce(W, s, wave_reduce_cache(W, tid), tid, nt);
            if (!wave_solve(W, s, *o, tick, limit, scaled)) {
                st = 2;
                it = 0;
                break;
            }
            int lim_any = 0;
            for (uint32_t tid = 0; tid < nt; tid++)
                for (uint32_t e = tid; e < wh.n_inst; e += nt) lim_any |= wave_limit_inst(W, s, e, lim_active) ? 1 : 0;
            lim_active = true;
            bool ok = wave_gates(W, s, *o, tick, lim_any != 0, limit, scaled);
            if (ok && (bits & G_INITFIX)) {
                bits &= ~G_INITFIX;
            } else if (ok) {
                for (uint32_t i = 0; i < n; i++) s.x[i] = s.x_old[i];
                st = 1;
                it = tick + 1;
                break;
            }
        }
        for (uint32_t i = 0; i < n; i++) x[(size_t)i * N + l] = s.x[i];
        status[l] = st;
        iters[l] = it;
    }
}

int hl_column_order(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *q)
{
    return column_order(n, col_ptr, row_idx, q);
}

int hl_amd(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *q)
{
    return amd_order(n, col_ptr, row_idx, q);
}

void hl_prep(int kind, const double *u, double *p) { prep_kind(kind, u, p); }

double hl_exp(double x) { return e_exp(x); }
double hl_log(double x) { return e_log(x); }
double hl_sin(double x) { return e_sin(x); }
double hl_pow(double x, double y) { return e_pow(x, y); }

}  // extern "C"

extern "C" {
// The workgroup-per-lane kernels' LDS bytes for this template: the static
// Newton's lane block and a transient's with n_passes .meas passes.
void hl_layout(void *tv, uint32_t n_passes, uint64_t *out)
{
    Template *t = (Template *)tv;
    WaveHost H;
    build_wave(*t, (uint32_t)std::max<size_t>((size_t)t->n * (t->n - 1), 1), H);
    const WaveTemplate W = wave_template(H.at, H.blob.data());
    WaveLayout L{}, LT{};
    // the lean layout of templates with compiled kinds, as the engine picks it
    bool va = false;
    for (int k = 0; k < K_NKIND; k++) va = va || (kind_is_va(k) && t->count[k] > 0);
    L.lu_global = LT.lu_global = va ? 1u : 0u;
    out[0] = 4 * (uint64_t)table_lds_words(H.at.words, L) + 8 * wave_words(W, L);
    out[1] = 4 * (uint64_t)table_lds_words(H.at.words, LT) + 8 * wave_words(W, LT, true, tran_extra_words(W.n_inst, n_passes));
    out[2] = 4 * (uint64_t)H.at.words;
    out[3] = H.n_prm;
    out[4] = H.n_prm_lds;
    out[5] = 8 * (uint64_t)(L.lu_u32 - L.lu_f64);
    out[6] = 8 * (uint64_t)(L.flags - L.lu_u32);
    out[7] = L.dense == ESP_NONE ? 0 : 8 * (uint64_t)dense_words(t->n, va);
    out[8] = 8 * (uint64_t)(L.prm - L.stage);
}
}
