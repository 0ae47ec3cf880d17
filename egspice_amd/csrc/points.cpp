// The points subset of include/espice.h on the lane engine.
//
// A Problem is one deck built into one lane template (deck.h). Its own
// analyses are point 0 (the deck's .param values); espice_append_points
// adds rows of live .param values. espice_run_all runs every pending point
// in batches of at most the lane cap, one GPU lane each: the deck's .op as
// a cold operating point (espice_lanes_newton_op) and its .tran as a whole
// device transient with the .meas cards as device passes
// (espice_lanes_tran). Query ids, handles, statuses and reasons follow
// src/c_api.zig and src/espice.zig.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/espice.h"
#include "../../include/espice_lanes.h"
#include "cli.h"
#include "deck.h"
#include "sampler.h"

using esp::BuiltDeck;

// rows of point 0's transient kept for the transient plot
constexpr uint32_t TRAN_PLOT_ROWS = 100000;
using esp::Meas;

namespace {

constexpr uint32_t LANE_CAP = 16384;  // the reference's gpu_lane_cap (docs/using/gpu.md:84-85)

struct Point {
    std::vector<double> live;  // live parameter values (point 0: the deck's)
    uint32_t first_id = 0;
    bool ran = false, released = false;
    uint32_t reason = ESPICE_REASON_OK;
    bool failed = false;       // no measure is read: NaN throughout
    std::vector<double> meas;  // one per .meas card
    std::vector<double> aux;   // two per card: the printed side values
    uint32_t tran_points = 0;
    uint32_t trial = 0;        // Monte Carlo trial (HSPICE numbering from 1), 0: nominal draws
    std::vector<double> x;     // the operating point (n rows), when one was solved
    // espice_points_options of the append that made the point
    // (include/espice.h:286-301): points run in batches of equal options
    struct Opts {
        bool early_stop = false, warm_start = false;
        double abort_after = NAN;
        int abort_card = -1;  // the card abort_measure names, -1 every card
        bool operator==(const Opts &o) const
        {
            return early_stop == o.early_stop && warm_start == o.warm_start && abort_card == o.abort_card &&
                   ((abort_after != abort_after && o.abort_after != o.abort_after) || abort_after == o.abort_after);
        }
    } opts;
};

// One .meas card as device passes.
struct CardPasses {
    int first = -1, second = -1;  // pass indices; -1 none
    uint32_t op_row = 0;          // `.meas op find`
};

struct Problem {
    BuiltDeck bd;
    espice_lanes_template *tpl = nullptr;
    std::vector<std::string> live;
    std::vector<Point> points;
    uint32_t per_point = 1;  // query ids per point
    uint32_t queries = 0;
    std::vector<espice_meas_pass> passes;
    std::vector<CardPasses> cards;
    std::vector<double> x_nominal;  // point 0's operating point: the start of `.dc monte` trials
    std::vector<double> x_tran_nominal;  // the template's nominal operating point (warm_start), NaN: none
    // point 0's `.ac` response at its operating point: the frequencies and
    // per frequency the real parts of every unknown, then the imaginary
    std::vector<double> ac_freq, ac_x;
    // point 0's transient rows when it ran alone: times, and every unknown
    // per row
    std::vector<double> tran_t, tran_x;
    // point 0's `.dc` sweep: the swept values and every unknown per value
    // (NaN where a value did not converge)
    std::vector<double> dc_v, dc_x;
    bool dc_is_v = true;
    uint32_t seed_trial = 0;        // the trial every non-Monte-Carlo point draws
    std::string error;
    ~Problem()
    {
        if (tpl) espice_lanes_template_destroy(tpl);
    }
};

// ------------------------------------------------------- handle table
// src/c_api.zig:185-268: a token is slot | generation << 12; slot 0 is never
// handed out; destroy bumps the generation, so stale tokens stop matching.
struct Handles {
    static constexpr unsigned BITS = 12, CAP = 1u << BITS;
    std::mutex mu;
    Problem *live[CAP] = {};
    uint64_t gen[CAP] = {};
    uint32_t free_list[CAP];
    uint32_t free_len = 0, fresh = 1;
};
Handles &handles()
{
    static Handles h;
    return h;
}
constexpr uint64_t GEN_MAX = (uint64_t(1) << (sizeof(uintptr_t) * 8 - Handles::BITS)) - 1;

espice_problem *insert(Problem *p)
{
    Handles &h = handles();
    std::lock_guard<std::mutex> g(h.mu);
    uint32_t idx;
    if (h.free_len) idx = h.free_list[--h.free_len];
    else if (h.fresh < Handles::CAP) idx = h.fresh++;
    else return nullptr;
    h.live[idx] = p;
    return reinterpret_cast<espice_problem *>((uintptr_t)(h.gen[idx] << Handles::BITS | idx));
}

int find_slot(Handles &h, espice_problem *tok)
{
    const uintptr_t b = reinterpret_cast<uintptr_t>(tok);
    const uint32_t idx = (uint32_t)(b & (Handles::CAP - 1));
    if (!h.live[idx] || h.gen[idx] != (uint64_t)(b >> Handles::BITS)) return -1;
    return (int)idx;
}

Problem *get(espice_problem *tok)
{
    Handles &h = handles();
    std::lock_guard<std::mutex> g(h.mu);
    const int i = find_slot(h, tok);
    return i < 0 ? nullptr : h.live[i];
}

Problem *remove(espice_problem *tok)
{
    Handles &h = handles();
    std::lock_guard<std::mutex> g(h.mu);
    const int i = find_slot(h, tok);
    if (i < 0) return nullptr;
    Problem *p = h.live[i];
    h.live[i] = nullptr;
    if (h.gen[i] == GEN_MAX) return p;  // retired
    h.gen[i]++;
    h.free_list[h.free_len++] = (uint32_t)i;
    return p;
}

espice_status fail(Problem *p, espice_status s, const std::string &e)
{
    if (p) p->error = e;
    return s;
}

std::string fold(const char *data, size_t len)
{
    std::string s(data ? data : "", data ? len : 0);
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

bool gpu_present()
{
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

// The row a vector names: v(node) or i(vsource).
bool vector_row(const BuiltDeck &bd, const std::string &vec, uint32_t &row)
{
    const std::string inner = vec.substr(2, vec.size() - 3);
    if (vec[0] == 'v') {
        if (inner == "0" || inner == "gnd") return false;
        auto it = bd.rows.find(inner);
        if (it == bd.rows.end()) return false;
        row = it->second;
        return true;
    }
    auto it = bd.vsrc_branch.find(inner);
    if (it == bd.vsrc_branch.end()) return false;
    row = it->second;
    return true;
}

double val_of(const esp::Value &v, const std::map<std::string, double> &env, double dflt, std::string &err)
{
    if (!v.given) return dflt;
    return v.expr ? esp::eval_expr(*v.expr, env, err) : v.num;
}

// The deck's .meas cards as device passes (their values are the deck's
// literals or .param expressions, fixed per problem: a pass's val/at/from
// may not depend on a live parameter).
bool plan_passes(Problem &P, std::string &err)
{
    std::map<std::string, double> env;
    if (!esp::param_env(P.bd.deck, {}, env, err)) return false;
    for (const Meas &m : P.bd.deck.meas) {
        CardPasses c;
        auto side_pass = [&](const Meas::Side &s, uint32_t kind, uint32_t op) -> int {
            espice_meas_pass q{};
            if (!vector_row(P.bd, s.vec, q.row)) {
                err = "NoSuchVector " + s.vec;
                return -1;
            }
            q.kind = kind;
            q.op = op;
            q.rise = s.rise;
            q.fall = s.fall;
            q.cross = s.cross;
            q.val = val_of(s.val, env, 1e99, err);
            q.td = val_of(s.td, env, 0.0, err);
            q.from = val_of(s.from, env, 0.0, err);
            q.to = val_of(s.to, env, 0.0, err);
            q.at = val_of(s.at, env, 0.0, err);
            if (!err.empty()) return -1;
            P.passes.push_back(q);
            return (int)P.passes.size() - 1;
        };
        if (m.analysis == "op") {
            if (m.kind != "find" || !vector_row(P.bd, m.a.vec, c.op_row)) {
                err = "UnsupportedMeasure " + m.name;
                return false;
            }
        } else if (m.kind == "trig_targ") {
            if ((c.first = side_pass(m.a, 0, 0)) < 0 || (c.second = side_pass(m.b, 0, 0)) < 0) return false;
        } else if (m.kind == "when") {
            if ((c.first = side_pass(m.a, 0, 0)) < 0) return false;
        } else if (m.kind == "find" || m.kind == "deriv") {
            if ((c.first = side_pass(m.a, 1, m.kind == "deriv" ? 1 : 0)) < 0) return false;
        } else if (m.kind == "find_when" || m.kind == "deriv_when") {
            // the WHEN's Cross, then the Find reading it; the card's value is the Find's
            if (side_pass(m.b, 0, 0) < 0 || (c.first = side_pass(m.a, 4, m.kind == "deriv_when" ? 1 : 0)) < 0)
                return false;
        } else if (m.kind == "min" || m.kind == "max" || m.kind == "avg") {
            if ((c.first = side_pass(m.a, 2, m.kind == "min" ? 0 : m.kind == "max" ? 1 : 2)) < 0) return false;
        } else if (m.kind == "rms" || m.kind == "integ") {
            if ((c.first = side_pass(m.a, 3, m.kind == "rms" ? 1 : 0)) < 0) return false;
        } else if (m.kind != "param") {
            err = "UnsupportedMeasure " + m.name;
            return false;
        }
        P.cards.push_back(c);
    }
    return true;
}

// Runs points [b, e) on one batch of lanes.
espice_status run_batch(Problem &P, size_t b, size_t e)
{
    const esp::Deck &d = P.bd.deck;
    const uint32_t N = (uint32_t)(e - b), n = P.bd.n;
    espice_lanes *L = nullptr;
    if (espice_lanes_create(P.tpl, N, &L) != ESPICE_OK) return fail(&P, ESPICE_FAILED, "GpuUnavailable");
    struct Guard {
        espice_lanes *l;
        ~Guard() { espice_lanes_destroy(l); }
    } guard{L};
    // per point .param environments, and every parameter slot per lane
    std::vector<std::map<std::string, double>> envs(N);
    std::vector<bool> bad(N, false);
    for (uint32_t l = 0; l < N; l++) {
        Point &pt = P.points[b + l];
        std::map<std::string, double> live;
        for (size_t k = 0; k < P.live.size(); k++) live[P.live[k]] = pt.live[k];
        std::string err;
        if (!esp::param_env(d, live, envs[l], err, pt.trial)) bad[l] = true;
    }
    std::vector<double> v(N);
    for (const esp::ParamSlot &s : P.bd.slots) {
        if (!s.value.expr) continue;
        for (uint32_t l = 0; l < N; l++) {
            std::string err;
            v[l] = bad[l] ? NAN : esp::slot_value(d, s, envs[l], err, P.points[b + l].trial);
            if (!err.empty() || !std::isfinite(v[l])) bad[l] = true;
        }
        if (espice_lanes_set_param(L, s.kind, s.inst, s.index, v.data()) != ESPICE_OK)
            return fail(&P, ESPICE_FAILED, "SetParam");
    }
    if (espice_lanes_prepare(L) != ESPICE_OK) return fail(&P, ESPICE_FAILED, "Prepare");
    const size_t nm = d.meas.size();
    for (uint32_t l = 0; l < N; l++) {
        Point &pt = P.points[b + l];
        pt.meas.assign(nm, NAN);
        pt.aux.assign(2 * nm, NAN);
        pt.reason = bad[l] ? ESPICE_REASON_INVALID_PARAMETER : ESPICE_REASON_OK;
        pt.failed = bad[l];
    }
    std::vector<double> xop;
    bool have_op = false;
    bool need_op = d.op;
    for (const Meas &m : d.meas) need_op = need_op || m.analysis == "op";
    if (d.op || d.monte_dc || d.ac_kind >= 0) {
        // a `.dc monte` trial starts from the nominal operating point with
        // no junction seeding (src/analysis/sweep/lanes.zig:73-83,
        // mc.zig:128-131); a batch holds trials only or no trials
        const bool warm = P.points[b].trial != 0 && d.monte_dc && !P.x_nominal.empty();
        std::vector<double> x0((size_t)n * N, 0.0);
        if (warm)
            for (uint32_t l = 0; l < N; l++)
                for (uint32_t i = 0; i < n; i++) x0[(size_t)i * N + l] = P.x_nominal[i];
        if (espice_lanes_upload_x(L, x0.data()) != ESPICE_OK) return fail(&P, ESPICE_FAILED, "Upload");
        // the plain rung on every lane, the stepping rungs for the lanes it
        // leaves (op.zig:626-688 solveGpuLanes)
        espice_newton_opts o{1e-3, 1e-12, 1e-6, 1e-9, 0.0, 100, warm ? 0u : 1u};
        const espice_ladder_opts lo{1e-2, 100, 0};
        std::vector<espice_lane_result> r(N);
        double ms;
        if (espice_lanes_op_ladder(L, &o, &lo, r.data(), nullptr, &ms) != ESPICE_OK)
            return fail(&P, ESPICE_FAILED, "NewtonOp");
        xop.resize((size_t)n * N);
        if (espice_lanes_download_x(L, xop.data()) != ESPICE_OK) return fail(&P, ESPICE_FAILED, "Download");
        have_op = true;
        for (uint32_t l = 0; l < N; l++) {
            Point &pt = P.points[b + l];
            if (pt.failed) continue;
            if (r[l].status != 1) {
                pt.failed = true;
                pt.reason = ESPICE_REASON_NO_CONVERGENCE;
                continue;
            }
            pt.x.resize(n);
            for (uint32_t i = 0; i < n; i++) pt.x[i] = xop[(size_t)i * N + l];
            if (b + l == 0) P.x_nominal = pt.x;
        }
        // `.ac` at point 0's operating point (src/analysis/ac/ac.zig:40-86):
        // the lanes' x is the ladder's answer
        if (d.ac_kind >= 0 && b == 0 && !P.points[0].failed) {
            std::vector<double> drive(2 * (size_t)n, 0.0);
            for (size_t k = 0; k < d.elements.size(); k++) {
                const esp::Element &e = d.elements[k];
                if (!e.ac || P.bd.elem_kind[k] < 0) continue;
                if (espice_lanes_template_ac_drive(P.tpl, (uint32_t)P.bd.elem_kind[k], P.bd.elem_inst[k], e.ac_mag,
                                                   e.ac_phase, drive.data()) != ESPICE_OK)
                    return fail(&P, ESPICE_FAILED, "AcDrive");
            }
            P.ac_freq = esp::ac_axis(d);
            std::vector<double> omega(P.ac_freq.size()), out(P.ac_freq.size() * 2 * n * N);
            for (size_t k = 0; k < omega.size(); k++) omega[k] = 2.0 * 3.14159265358979323846 * P.ac_freq[k];
            double ms;
            if (espice_lanes_ac(L, omega.data(), (uint32_t)omega.size(), drive.data(), out.data(), &ms) != ESPICE_OK) {
                size_t req;
                char buf[128];
                espice_lanes_error_message(buf, sizeof buf, &req);
                return fail(&P, ESPICE_FAILED, std::string("Ac: ") + buf);
            }
            P.ac_x.assign(omega.size() * 2 * n, 0.0);
            for (size_t k = 0; k < omega.size(); k++)
                for (uint32_t part = 0; part < 2; part++)
                    for (uint32_t i = 0; i < n; i++)
                        P.ac_x[(k * 2 + part) * n + i] = out[((k * 2 + part) * n + i) * N];
        }
    }
    if (d.tran) {
        std::map<std::string, double> env;
        std::string err;
        esp::param_env(d, {}, env, err);
        espice_tran_opts o{};
        o.t_stop = val_of(d.tstop, env, 0, err);
        o.dt_init = val_of(d.tstep, env, 0, err);
        o.dt_max = d.tmax.given ? val_of(d.tmax, env, 0, err) : std::fmin(o.dt_init, o.t_stop / 50);
        o.dt_min = 1e-18;
        o.xmu = 0.5;
        o.reltol = 1e-3;
        o.abstol = 1e-12;
        o.vntol = 1e-6;
        o.chgtol = 1e-14;
        o.trtol = 7.0;
        o.method = 1;
        o.max_iter = 100;
        o.max_steps = 1000000000u;
        o.record_capacity = need_op && !have_op ? 1 : 0;
        // point 0 alone in its batch (the command line's run): its waveform
        // is kept for the transient plot, up to TRAN_PLOT_ROWS rows
        const bool plot = b == 0 && N == 1;
        if (plot) o.record_capacity = TRAN_PLOT_ROWS;
        // the batch's points options (one set per batch)
        const Point::Opts &po = P.points[b].opts;
        o.early_stop = po.early_stop ? 1u : 0u;
        o.abort_after = po.abort_after;
        o.abort_mask = 0;
        for (size_t k = 0; k < d.meas.size(); k++) {
            if (po.abort_card >= 0 && (int)k != po.abort_card) continue;
            const CardPasses &c = P.cards[k];
            for (int q : {c.first, c.second})
                if (q >= 0 && q < 64) o.abort_mask |= 1ull << q;
            // a FIND ... WHEN card's WHEN is the pass before its Find
            if ((d.meas[k].kind == "find_when" || d.meas[k].kind == "deriv_when") && c.first > 0 && c.first - 1 < 64)
                o.abort_mask |= 1ull << (c.first - 1);
        }
        // warm_start: every point's operating point from the template's
        // nominal one (src/analysis/dc/op.zig:588, 671), solved cold once by
        // the ladder
        if (po.warm_start) {
            if (P.x_tran_nominal.empty()) {
                espice_lanes *L0 = nullptr;
                if (espice_lanes_create(P.tpl, 1, &L0) != ESPICE_OK) return fail(&P, ESPICE_FAILED, "GpuUnavailable");
                std::vector<double> z(n, 0.0), x1(n);
                espice_newton_opts no{1e-3, 1e-12, 1e-6, 1e-9, 0.0, 100, 1u};
                const espice_ladder_opts lo{1e-2, 100, 0};
                espice_lane_result r1{};
                double ms0;
                bool ok = espice_lanes_prepare(L0) == ESPICE_OK && espice_lanes_upload_x(L0, z.data()) == ESPICE_OK &&
                          espice_lanes_op_ladder(L0, &no, &lo, &r1, nullptr, &ms0) == ESPICE_OK && r1.status == 1 &&
                          espice_lanes_download_x(L0, x1.data()) == ESPICE_OK;
                espice_lanes_destroy(L0);
                P.x_tran_nominal = ok ? x1 : std::vector<double>(n, NAN);
            }
            if (std::isfinite(P.x_tran_nominal[0])) {
                std::vector<double> xw((size_t)n * N);
                for (uint32_t l = 0; l < N; l++)
                    for (uint32_t i = 0; i < n; i++) xw[(size_t)i * N + l] = P.x_tran_nominal[i];
                if (espice_lanes_upload_x(L, xw.data()) != ESPICE_OK) return fail(&P, ESPICE_FAILED, "Upload");
                o.op_warm = 1;
            }
        }
        std::vector<espice_tran_result> tr(N);
        std::vector<espice_meas_result> mr(P.passes.size() * N + 1);
        const size_t cap = o.record_capacity;
        std::vector<double> t0(cap ? N * cap : 1), x0(cap ? (size_t)N * cap * n : 1);
        double ms;
        const espice_status st = espice_lanes_tran(L, &o, P.passes.data(), (uint32_t)P.passes.size(), tr.data(),
                                                   mr.data(), t0.data(), x0.data(), &ms);
        if (st != ESPICE_OK) {
            size_t req;
            char buf[128];
            espice_lanes_error_message(buf, sizeof buf, &req);
            return fail(&P, st, std::string("Tran: ") + buf);
        }
        if (plot) {
            const size_t rows = std::min<size_t>(tr[0].points, cap);
            P.tran_t.assign(t0.begin(), t0.begin() + rows);
            P.tran_x.assign(x0.begin(), x0.begin() + rows * n);
        }
        if (cap && !have_op && need_op) {
            // the transient's operating point: row t = 0, laid out like x
            xop.assign((size_t)n * N, NAN);
            for (uint32_t l = 0; l < N; l++)
                for (uint32_t i = 0; i < n; i++) xop[(size_t)i * N + l] = x0[(size_t)l * cap * n + i];
            have_op = true;
        }
        for (uint32_t l = 0; l < N; l++) {
            Point &pt = P.points[b + l];
            if (pt.failed) continue;
            const uint32_t s = tr[l].status;
            pt.tran_points = tr[l].points;
            if (s != 0) {
                pt.failed = true;
                pt.reason = s == 1   ? ESPICE_REASON_TIMESTEP_TOO_SMALL
                            : s == 3 ? ESPICE_REASON_SINGULAR_MATRIX
                            : s == 4 ? ESPICE_REASON_ABORTED
                                     : ESPICE_REASON_NO_CONVERGENCE;
                continue;
            }
            const espice_meas_result *m = &mr[(size_t)l * P.passes.size()];
            for (size_t k = 0; k < nm; k++) {
                const Meas &card = d.meas[k];
                const CardPasses &c = P.cards[k];
                if (card.analysis != "tran" || card.kind == "param") continue;
                const espice_meas_result &a = m[c.first];
                if (card.kind == "trig_targ") {
                    const double trig = a.value, targ = m[c.second].value;
                    pt.meas[k] = targ - trig;
                    pt.aux[2 * k] = targ;
                    pt.aux[2 * k + 1] = trig;
                } else {
                    pt.meas[k] = a.value;
                    pt.aux[2 * k] = a.aux0;
                    pt.aux[2 * k + 1] = a.aux1;
                }
            }
        }
    }
    // `.dc` last, as it moves its source (dc.zig:161-293 runGpuLanes): every
    // lane's sweep, point 0's kept for the DC plot
    if (!d.dc_src.empty() && b == 0) {
        int kind = -1;
        uint32_t inst = 0;
        for (size_t k = 0; k < d.elements.size(); k++)
            if (d.elements[k].name == d.dc_src && P.bd.elem_kind[k] >= 0) {
                kind = P.bd.elem_kind[k];
                inst = P.bd.elem_inst[k];
            }
        if (kind < 0) return fail(&P, ESPICE_INVALID_ARGUMENT, "DcSweepSourceNotFound " + d.dc_src);
        const std::vector<double> v = esp::dc_values(d);
        espice_newton_opts o{1e-3, 1e-12, 1e-6, 1e-9, 0.0, 100, 1u};
        const espice_ladder_opts lo{1e-2, 100, 0};
        std::vector<double> xs(v.size() * (size_t)n * N);
        std::vector<uint32_t> conv(v.size() * (size_t)N);
        double ms;
        if (espice_lanes_dc_sweep(L, (uint32_t)kind, inst, v.data(), (uint32_t)v.size(), &o, &lo, xs.data(),
                                  conv.data(), &ms) != ESPICE_OK) {
            size_t req;
            char buf[128];
            espice_lanes_error_message(buf, sizeof buf, &req);
            return fail(&P, ESPICE_FAILED, std::string("Dc: ") + buf);
        }
        P.dc_v = v;
        P.dc_x.assign(v.size() * (size_t)n, NAN);
        for (size_t k = 0; k < v.size(); k++)
            for (uint32_t i = 0; i < n; i++) P.dc_x[k * n + i] = xs[(k * n + i) * N];
        P.dc_is_v = d.dc_src[0] == 'v';
    }
    for (uint32_t l = 0; l < N; l++) {
        Point &pt = P.points[b + l];
        if (pt.failed) {
            pt.meas.assign(nm, NAN);
            pt.ran = true;
            continue;
        }
        // .meas op, then PARAM cards over the other cards' results (measure.c
        // order: PARAM cards read every other card)
        std::map<std::string, double> env = envs[l];
        for (size_t k = 0; k < nm; k++) {
            const Meas &card = d.meas[k];
            if (card.analysis == "op" && have_op) pt.meas[k] = xop[(size_t)P.cards[k].op_row * N + l];
            if (card.kind != "param") env[card.name] = pt.meas[k];
        }
        for (size_t k = 0; k < nm; k++) {
            const Meas &card = d.meas[k];
            if (card.kind != "param") continue;
            std::string err;
            pt.meas[k] = esp::eval_expr(*card.param, env, err);
            if (!err.empty()) pt.meas[k] = NAN;
            env[card.name] = pt.meas[k];
        }
        for (size_t k = 0; k < nm; k++)
            if (std::isnan(pt.meas[k]) && pt.reason == ESPICE_REASON_OK) pt.reason = ESPICE_REASON_MEASURE_UNRESOLVED;
        pt.ran = true;
    }
    return ESPICE_OK;
}

// The point whose first query is `id`.
int point_at(const Problem &P, uint32_t id)
{
    for (size_t i = 0; i < P.points.size(); i++)
        if (P.points[i].first_id == id) return (int)i;
    return -1;
}

}  // namespace

extern "C" {

uint32_t espice_abi_version(void) { return ESPICE_ABI_VERSION; }

void espice_default_options(espice_create_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->abi_version = ESPICE_ABI_VERSION;
    o->struct_size = sizeof *o;
    o->max_parallel = 1;
}

espice_status espice_create(const espice_create_options *o, espice_problem **out, char *diag, size_t cap)
{
    auto say = [&](espice_status s, const std::string &e) {
        if (diag && cap) {
            const size_t k = e.size() < cap - 1 ? e.size() : cap - 1;
            std::memcpy(diag, e.data(), k);
            diag[k] = 0;
        }
        return s;
    };
    if (out) *out = nullptr;
    if (!o || !out) return say(ESPICE_INVALID_ARGUMENT, "NullArgument");
    if (o->abi_version != ESPICE_ABI_VERSION || o->struct_size < offsetof(espice_create_options, zig_path))
        return say(ESPICE_ABI_MISMATCH, "AbiMismatch");
    if (o->source_kind > ESPICE_BYTES || o->dialect > ESPICE_SPECTRE || o->backend > ESPICE_HIP ||
        o->explicit_gpu > 1 || o->max_parallel == 0 || o->max_parallel > 65535 || !o->source.data || !o->source.len)
        return say(ESPICE_INVALID_ARGUMENT, "InvalidArgument");
    if (o->dialect != ESPICE_NGSPICE) return say(ESPICE_INVALID_ARGUMENT, "UnsupportedDialect");
    if (o->backend == ESPICE_CPU || o->backend == ESPICE_CUDA) return say(ESPICE_FAILED, "BackendUnavailable");
    std::string text;
    if (o->source_kind == ESPICE_FILE) {
        std::ifstream f(std::string(o->source.data, o->source.len), std::ios::binary);
        if (!f) return say(ESPICE_FAILED, "FileNotFound");
        std::stringstream ss;
        ss << f.rdbuf();
        text = ss.str();
    } else {
        text.assign(o->source.data, o->source.len);
    }
    auto P = std::make_unique<Problem>();
    std::string err;
    if (!esp::read_deck(text, P->bd.deck, err)) return say(ESPICE_INVALID_ARGUMENT, err);
    const esp::Deck &d = P->bd.deck;
    if (!d.op && !d.tran && !d.monte_dc && d.dc_src.empty() && d.ac_kind < 0)
        return say(ESPICE_INVALID_ARGUMENT, "NoAnalysis");
    if (espice_lanes_template_create(&P->tpl) != ESPICE_OK) return say(ESPICE_OUT_OF_MEMORY, "OutOfMemory");
    if (!esp::build_deck(P->bd, P->tpl, err)) return say(ESPICE_INVALID_ARGUMENT, err);
    for (size_t k = 0; k < o->live_param_count; k++) {
        const std::string name = fold(o->live_params[k].data, o->live_params[k].len);
        bool found = false;
        for (auto &pr : d.params) found = found || pr.first == name;
        if (!found) return say(ESPICE_INVALID_ARGUMENT, "NotAParameter " + name);
        P->live.push_back(name);
    }
    if (!plan_passes(*P, err)) return say(ESPICE_INVALID_ARGUMENT, err);
    if (!gpu_present()) return say(ESPICE_FAILED, "GpuUnavailable");
    // point 0: the deck as written
    Point p0;
    std::map<std::string, double> env;
    esp::param_env(d, {}, env, err);
    for (auto &name : P->live) p0.live.push_back(env[name]);
    // `.option seed` without a Monte Carlo card draws trial 1 for every point
    // (the deck's mismatch sample, src/frontend/netlist.zig:1401,
    // variants.zig:1088-1090, 1172-1176)
    P->seed_trial = d.has_seed && !d.monte_dc && !d.monte_tran ? 1 : 0;
    p0.trial = P->seed_trial;
    P->per_point = (d.op || d.monte_dc ? 1u : 0u) + (d.tran ? 1u : 0u);
    p0.first_id = 0;
    P->queries = P->per_point;
    P->points.push_back(p0);
    // `.dc monte=N` or `.tran ... sweep monte=N`: trials 1..N, each a point
    // of its own after the nominal
    for (uint32_t k = 1; k <= d.monte_dc + d.monte_tran; k++) {
        Point pt;
        pt.live = p0.live;
        pt.trial = k;
        pt.first_id = P->queries;
        P->queries += P->per_point;
        P->points.push_back(std::move(pt));
    }
    espice_problem *tok = insert(P.get());
    if (!tok) return say(ESPICE_OUT_OF_MEMORY, "TooManyProblems");
    P.release();
    *out = tok;
    return ESPICE_OK;
}

espice_status espice_destroy(espice_problem *problem)
{
    if (!problem) return ESPICE_OK;
    Problem *p = remove(problem);
    if (!p) return ESPICE_INVALID_HANDLE;
    delete p;
    return ESPICE_OK;
}

espice_status espice_query_count(espice_problem *problem, uint32_t *out)
{
    if (!problem) return ESPICE_INVALID_ARGUMENT;
    Problem *P = get(problem);
    if (!P) return ESPICE_INVALID_HANDLE;
    if (!out) return fail(P, ESPICE_INVALID_ARGUMENT, "NullArgument");
    *out = P->queries;
    return ESPICE_OK;
}

espice_status espice_error_message(espice_problem *problem, char *buf, size_t cap, size_t *required)
{
    if (!problem) return ESPICE_INVALID_ARGUMENT;
    Problem *P = get(problem);
    if (!P) return ESPICE_INVALID_HANDLE;
    const std::string &e = P->error;
    if (required) *required = e.size() + 1;
    if (cap < e.size() + 1) return ESPICE_BUFFER_TOO_SMALL;
    std::memcpy(buf, e.c_str(), e.size() + 1);
    return ESPICE_OK;
}

espice_status espice_run_all(espice_problem *problem)
{
    if (!problem) return ESPICE_INVALID_ARGUMENT;
    Problem *P = get(problem);
    if (!P) return ESPICE_INVALID_HANDLE;
    size_t b = 0;
    const size_t n = P->points.size();
    // the nominal point first and alone when trials start from its
    // operating point
    if (P->bd.deck.monte_dc && !P->points[0].ran) {
        const espice_status s = run_batch(*P, 0, 1);
        if (s != ESPICE_OK) return s;
    }
    while (b < n) {
        while (b < n && (P->points[b].ran || P->points[b].released)) b++;
        if (b >= n) break;
        size_t e = b;
        const bool trial = P->points[b].trial != 0 && P->bd.deck.monte_dc;
        while (e < n && !P->points[e].ran && !P->points[e].released && e - b < LANE_CAP &&
               (P->points[e].trial != 0 && P->bd.deck.monte_dc) == trial && P->points[e].opts == P->points[b].opts)
            e++;
        const espice_status s = run_batch(*P, b, e);
        if (s != ESPICE_OK) return s;
        b = e;
    }
    return ESPICE_OK;
}

espice_status espice_append_points(espice_problem *problem, const double *values, size_t count,
                                   const espice_points_options *opts, uint32_t *first_id)
{
    if (!problem) return ESPICE_INVALID_ARGUMENT;
    Problem *P = get(problem);
    if (!P) return ESPICE_INVALID_HANDLE;
    if (opts && opts->struct_size < offsetof(espice_points_options, tangent_step))
        return fail(P, ESPICE_ABI_MISMATCH, "AbiMismatch");
    if (count == 0 || !values || !first_id) return fail(P, ESPICE_INVALID_ARGUMENT, "InvalidArgument");
    if (P->live.empty()) return fail(P, ESPICE_INVALID_ARGUMENT, "NoLiveParameters");
    Point::Opts po;
    if (opts) {
        if (opts->measures_only > 1 || opts->early_stop > 1 || opts->warm_start > 1)
            return fail(P, ESPICE_INVALID_ARGUMENT, "InvalidArgument");
        if (!std::isnan(opts->abort_after) && (opts->abort_after < 0 || std::isinf(opts->abort_after)))
            return fail(P, ESPICE_INVALID_ARGUMENT, "InvalidArgument");
        if (opts->abort_measure.len) {
            const std::string name = fold(opts->abort_measure.data, opts->abort_measure.len);
            bool found = false;
            for (size_t k = 0; k < P->bd.deck.meas.size(); k++)
                if (P->bd.deck.meas[k].name == name) {
                    found = true;
                    po.abort_card = (int)k;
                }
            if (!found) return fail(P, ESPICE_INVALID_ARGUMENT, "UnknownMeasure");
        }
        if (opts->struct_size >= sizeof(espice_points_options) && opts->tangent_step > 0)
            return fail(P, ESPICE_INVALID_ARGUMENT, "TangentsUnavailable");
        // measures_only: the lanes measure as the .tran runs and keep no
        // waveform in any case, so it changes nothing here
        po.early_stop = opts->early_stop != 0;
        po.warm_start = opts->warm_start != 0;
        po.abort_after = opts->abort_after;
    }
    const size_t k = P->live.size();
    *first_id = P->queries;
    for (size_t i = 0; i < count; i++) {
        Point pt;
        pt.live.assign(values + i * k, values + (i + 1) * k);
        pt.opts = po;
        pt.trial = P->seed_trial;
        pt.first_id = P->queries;
        P->queries += P->per_point;
        bool finite = true;
        for (double v : pt.live) finite = finite && std::isfinite(v);
        if (!finite) {
            pt.ran = true;
            pt.failed = true;
            pt.reason = ESPICE_REASON_INVALID_PARAMETER;
            pt.meas.assign(P->bd.deck.meas.size(), NAN);
        }
        P->points.push_back(std::move(pt));
    }
    return ESPICE_OK;
}

espice_status espice_copy_point_measures(espice_problem *problem, uint32_t first_id, size_t count,
                                         const espice_bytes *names, size_t measure_count, double *out,
                                         uint32_t *reasons)
{
    if (!problem) return ESPICE_INVALID_ARGUMENT;
    Problem *P = get(problem);
    if (!P) return ESPICE_INVALID_HANDLE;
    if ((measure_count && (!names || !out)) || (count && !reasons))
        return fail(P, ESPICE_INVALID_ARGUMENT, "NullArgument");
    std::vector<size_t> idx;
    for (size_t j = 0; j < measure_count; j++) {
        const std::string name = fold(names[j].data, names[j].len);
        size_t k = 0;
        while (k < P->bd.deck.meas.size() && P->bd.deck.meas[k].name != name) k++;
        if (k == P->bd.deck.meas.size()) return fail(P, ESPICE_INVALID_ARGUMENT, "UnknownMeasure");
        idx.push_back(k);
    }
    const int p0 = point_at(*P, first_id);
    if (p0 < 0 || (size_t)p0 + count > P->points.size()) return fail(P, ESPICE_INVALID_QUERY, "InvalidQuery");
    for (size_t i = 0; i < count; i++) {
        const Point &pt = P->points[p0 + i];
        if (!pt.ran || pt.released) return fail(P, ESPICE_RESULT_UNAVAILABLE, "ResultUnavailable");
    }
    for (size_t i = 0; i < count; i++) {
        const Point &pt = P->points[p0 + i];
        for (size_t j = 0; j < measure_count; j++) out[i * measure_count + j] = pt.failed ? NAN : pt.meas[idx[j]];
        reasons[i] = pt.reason;
    }
    return ESPICE_OK;
}

espice_status espice_release_points(espice_problem *problem, uint32_t first_id, size_t count)
{
    if (!problem) return ESPICE_INVALID_ARGUMENT;
    Problem *P = get(problem);
    if (!P) return ESPICE_INVALID_HANDLE;
    const int p0 = point_at(*P, first_id);
    if (p0 < 0 || (size_t)p0 + count > P->points.size()) return fail(P, ESPICE_INVALID_QUERY, "InvalidQuery");
    for (size_t i = 0; i < count; i++) {
        const Point &pt = P->points[p0 + i];
        if (!pt.ran && !pt.released) return fail(P, ESPICE_RESULT_UNAVAILABLE, "ResultUnavailable");
    }
    for (size_t i = 0; i < count; i++) {
        Point &pt = P->points[p0 + i];
        pt.released = true;
        std::vector<double>().swap(pt.meas);
        std::vector<double>().swap(pt.aux);
    }
    return ESPICE_OK;
}

espice_status espice_lane_width(espice_problem *problem, uint32_t *out)
{
    if (!problem) return ESPICE_INVALID_ARGUMENT;
    Problem *P = get(problem);
    if (!P) return ESPICE_INVALID_HANDLE;
    if (!out) return fail(P, ESPICE_INVALID_ARGUMENT, "NullArgument");
    *out = LANE_CAP;
    return ESPICE_OK;
}

}  // extern "C"

// ---- the CLI's view of point 0 (not part of espice.h): the printed lines
// of the deck's .meas cards in the reference's format
// (src/output/measure.zig:220-237, 948-1051), then its summary.
namespace esp {

// C's %e with `prec` digits (glibc's rounding of the exact decimal).
static std::string sci(double x, int prec)
{
    char b[64];
    std::snprintf(b, sizeof b, "%.*e", prec, x);
    return b;
}

std::string problem_report(espice_problem *problem)
{
    Problem *P = get(problem);
    if (!P || P->points.empty() || !P->points[0].ran) return "";
    const Point &pt = P->points[0];
    const esp::Deck &d = P->bd.deck;
    std::string out;
    auto head = [&](const char *what) { out += std::string("\n  Measurements for ") + what + " Analysis\n\n"; };
    auto line = [&](const std::string &name, const std::string &body) {
        char b[32];
        std::snprintf(b, sizeof b, "%-20s", name.c_str());
        out += std::string(b) + "=" + body + "\n";
    };
    // a card's printed line for one point
    auto card_line = [&](const Point &pt, size_t k) {
        const Meas &m = d.meas[k];
        const double v = pt.failed ? NAN : pt.meas[k];
        const double a0 = pt.aux.size() > 2 * k ? pt.aux[2 * k] : NAN;
        const double a1 = pt.aux.size() > 2 * k ? pt.aux[2 * k + 1] : NAN;
        if (std::isnan(v)) {
            line(m.name, "  NaN");
            return;
        }
        if (m.kind == "trig_targ") line(m.name, "  " + sci(v, 6) + " targ=  " + sci(a0, 6) + " trig=  " + sci(a1, 6));
        else if (m.kind == "when") line(m.name, "   " + sci(v, 5));
        else if (m.kind == "rms" || m.kind == "integ")
            line(m.name, "   " + sci(v, 5) + " from=  " + sci(a0, 5) + " to=  " + sci(a1, 5));
        else if (m.kind == "avg") {
            std::map<std::string, double> env;
            std::string err;
            esp::param_env(d, {}, env, err);
            const double from = val_of(m.a.from, env, 0.0, err);
            line(m.name, "  " + sci(v, 6) + " from=  " + sci(from, 6) + " to=  " + sci(a0, 6));
        } else if (m.kind == "min" || m.kind == "max")
            line(m.name, "  " + sci(v, 6) + " at=  " + sci(a0, 6));
        else if (m.kind == "param") line(m.name, "  " + sci(v, 5));
        else line(m.name, "  " + sci(v, 6));
    };
    for (const char *an : {"op", "tran"}) {
        bool headed = false;
        for (size_t k = 0; k < d.meas.size(); k++) {
            if (d.meas[k].analysis != an) continue;
            if (!headed) head(std::string(an) == "op" ? "Operating Point" : "Transient");
            headed = true;
            card_line(pt, k);
        }
    }
    if (d.monte_tran) {
        // each trial's lines, then every card's statistics over the trials
        // where it is finite: Welford's mean and M2, sigma with Bessel's
        // correction (src/output/measure.zig:816-830, 904-943)
        std::vector<const Point *> trials;
        for (const Point &q : P->points)
            if (q.trial != 0 && q.ran) trials.push_back(&q);
        for (const Point *q : trials)
            for (size_t k = 0; k < d.meas.size(); k++)
                if (d.meas[k].analysis == "tran") card_line(*q, k);
        bool heading = false;
        for (size_t k = 0; k < d.meas.size(); k++) {
            if (d.meas[k].analysis != "tran") continue;
            double n = 0, mean = 0, m2 = 0, lo = INFINITY, hi = -INFINITY;
            for (const Point *q : trials) {
                const double x = q->failed ? NAN : q->meas[k];
                if (!std::isfinite(x)) continue;
                n += 1;
                const double delta = x - mean;
                mean += delta / n;
                m2 += delta * (x - mean);
                lo = std::fmin(lo, x);
                hi = std::fmax(hi, x);
            }
            if (n == 0) continue;
            if (!heading) out += "\n  Monte Carlo statistics over " + std::to_string(trials.size()) + " trials\n\n";
            heading = true;
            const std::string &nm = d.meas[k].name;
            const double sigma = n > 1 ? std::sqrt(m2 / (n - 1)) : 0;
            out += "mean(" + nm + ") = " + sci(mean, 6) + "\nsigma(" + nm + ") = " + sci(sigma, 6) + "\nmin(" + nm +
                   ") = " + sci(lo, 6) + "\nmax(" + nm + ") = " + sci(hi, 6) + "\n";
        }
    }
    out += d.title + ": " + std::to_string(d.elements.size()) + " devices\n";
    if (d.op) out += "  Operating Point: " + std::string(pt.failed ? "failed" : "done") + "\n";
    if (d.tran)
        // the vectors are time, every named node and every V source's
        // branch current; device-internal rows are not output variables
        out += "  Transient Analysis: " + std::to_string(pt.tran_points) + " points, " +
               std::to_string(1 + P->bd.rows.size() + P->bd.vsrc_branch.size()) + " variables\n";
    for (const std::string &w : d.ignored) out += "note: " + w + " is not run by the lane engine\n";
    return out;
}

// f64 as the reference's `{e}` prints it: the shortest digits that read
// back to the same double, then `e` and the exponent without sign padding
// (`1e0`, `-4.98898735103262e-4`, `5e-1`).
static std::string shortest_e(double x)
{
    if (std::isnan(x)) return "nan";
    if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
    if (x == 0) return std::signbit(x) ? "-0e0" : "0e0";
    char b[64];
    for (int p = 0; p <= 17; p++) {
        std::snprintf(b, sizeof b, "%.*e", p, x);
        if (std::strtod(b, nullptr) == x) break;
    }
    const std::string t(b);
    const size_t e = t.find('e');
    return t.substr(0, e) + "e" + std::to_string(std::atoi(t.c_str() + e + 1));
}

// One plot of the run, as the reference's output writers take it
// (src/output/types.zig Plot): variable names, real or complex, and the
// points row by row (a complex variable's real then imaginary part).
struct Plot {
    std::string plotname;
    std::vector<std::string> names;
    bool complex = false;
    size_t npoints = 0;
    std::vector<double> data;
};

// The run's plots: point 0's AC response (src/analysis/ac/ac.zig:56-85:
// frequency, then every probe as the builder publishes them, V-source
// currents in card order and then every named node in row order,
// frontend/builder.zig:740-760), and for `.dc monte=N` the "Monte Carlo"
// plot, one row per converged trial: its ordinal, the trial number as
// `run`, then every node voltage and V-source current, names in order
// (src/analysis/sweep/mc.zig:148-171).
// The builder's probes (frontend/builder.zig:740-760): V-source branch
// currents in card order, then every named node in row order.
static std::vector<std::pair<std::string, uint32_t>> builder_probes(Problem *P)
{
    const esp::Deck &d = P->bd.deck;
    std::vector<std::pair<std::string, uint32_t>> probes;
    for (size_t k = 0; k < d.elements.size(); k++) {
        const esp::Element &e = d.elements[k];
        if (e.letter != 'v') continue;
        auto it = P->bd.vsrc_branch.find(e.name);
        if (it != P->bd.vsrc_branch.end()) probes.emplace_back("i(" + e.name + ")", it->second);
    }
    std::vector<std::pair<uint32_t, std::string>> nodes;
    for (auto &[name, row] : P->bd.rows) nodes.emplace_back(row, name);
    std::sort(nodes.begin(), nodes.end());
    for (auto &[row, name] : nodes) probes.emplace_back("v(" + name + ")", row);
    return probes;
}

std::vector<Plot> problem_plots(Problem *P)
{
    std::vector<Plot> plots;
    const esp::Deck &d = P->bd.deck;
    if (d.ac_kind >= 0 && !P->ac_freq.empty()) {
        const uint32_t n = P->bd.n;
        const auto probes = builder_probes(P);
        Plot p;
        p.plotname = "AC Analysis";
        p.complex = true;
        p.names.push_back("frequency");
        for (auto &pr : probes) p.names.push_back(pr.first);
        p.npoints = P->ac_freq.size();
        for (size_t k = 0; k < p.npoints; k++) {
            p.data.push_back(P->ac_freq[k]);
            p.data.push_back(0.0);
            for (auto &pr : probes) {
                p.data.push_back(P->ac_x[(k * 2) * n + pr.second]);
                p.data.push_back(P->ac_x[(k * 2 + 1) * n + pr.second]);
            }
        }
        plots.push_back(std::move(p));
    }
    if (!d.dc_src.empty() && !P->dc_v.empty()) {
        // "DC transfer characteristic": the sweep column named after the
        // swept quantity (dc.zig:534-549 sweepColumn), then the probes
        const uint32_t n = P->bd.n;
        Plot p;
        p.plotname = "DC transfer characteristic";
        p.names.push_back(P->dc_is_v ? "v(v-sweep)" : "i(i-sweep)");
        const auto probes = builder_probes(P);
        for (auto &pr : probes) p.names.push_back(pr.first);
        p.npoints = P->dc_v.size();
        for (size_t k = 0; k < p.npoints; k++) {
            p.data.push_back(P->dc_v[k]);
            for (auto &pr : probes) p.data.push_back(P->dc_x[k * n + pr.second]);
        }
        plots.push_back(std::move(p));
    }
    if (d.tran && !P->tran_t.empty()) {
        const uint32_t n = P->bd.n;
        Plot p;
        p.plotname = "Transient Analysis";
        p.names.push_back("time");
        const auto probes = builder_probes(P);
        for (auto &pr : probes) p.names.push_back(pr.first);
        p.npoints = P->tran_t.size();
        for (size_t k = 0; k < p.npoints; k++) {
            p.data.push_back(P->tran_t[k]);
            for (auto &pr : probes) p.data.push_back(P->tran_x[k * n + pr.second]);
        }
        plots.push_back(std::move(p));
    }
    if (d.monte_dc) {
        std::vector<std::pair<std::string, uint32_t>> probes;
        for (auto &[name, row] : P->bd.rows) probes.emplace_back("v(" + name + ")", row);
        for (auto &[name, row] : P->bd.vsrc_branch) probes.emplace_back("i(" + name + ")", row);
        std::sort(probes.begin(), probes.end());
        Plot p;
        p.plotname = "Monte Carlo";
        p.names.push_back("run");
        for (auto &pr : probes) p.names.push_back(pr.first);
        for (const Point &pt : P->points) {
            if (pt.trial == 0 || !pt.ran || pt.failed || pt.x.empty()) continue;
            p.data.push_back((double)pt.trial);
            for (auto &pr : probes) p.data.push_back(pt.x[pr.second]);
            p.npoints++;
        }
        plots.push_back(std::move(p));
    }
    return plots;
}

// ngspice's variable type of a name (src/output/rawfile.zig:10-16).
static const char *var_type(const std::string &name)
{
    if (name == "time") return "time";
    if (name == "frequency") return "frequency";
    if (name.rfind("i(", 0) == 0 || name.find("#branch") != std::string::npos) return "current";
    return "voltage";
}

// The plots written as one of the reference's formats: 0 SPICE3 print
// (src/output/spice_print.zig:10-44), 1 CSV (csv.zig: a header row, a
// complex variable as _re and _im columns, RFC 4180 quoting), 2 ASCII raw
// (ascii_raw.zig), 3 binary raw (rawfile.zig: the text header with the
// fixed date, then the samples as native-endian f64).
std::string encode_plots(const std::vector<Plot> &plots, const std::string &title, int format)
{
    std::string out;
    for (const Plot &p : plots) {
        const size_t stride = p.names.size() * (p.complex ? 2 : 1);
        if (format == 0) {
            out += p.plotname + ": " + title + "\nIndex";
            for (auto &nm : p.names) out += "\t" + nm + (p.complex ? "\t(imag)" : "");
            out += "\n-----";
            for (size_t k = 0; k < stride; k++) out += "\t---------------";
            out += "\n";
            for (size_t k = 0; k < p.npoints; k++) {
                out += std::to_string(k);
                for (size_t v = 0; v < stride; v++) out += "\t" + shortest_e(p.data[k * stride + v]);
                out += "\n";
            }
        } else if (format == 1) {
            auto field = [&](const std::string &nm, const char *suffix) {
                if (nm.find_first_of(",\"\r\n") == std::string::npos) {
                    out += nm + suffix;
                    return;
                }
                out += "\"";
                for (char ch : nm) {
                    if (ch == '"') out += "\"";
                    out += ch;
                }
                out += std::string(suffix) + "\"";
            };
            for (size_t i = 0; i < p.names.size(); i++) {
                if (i) out += ",";
                if (p.complex) {
                    field(p.names[i], "_re");
                    out += ",";
                    field(p.names[i], "_im");
                } else {
                    field(p.names[i], "");
                }
            }
            out += "\n";
            for (size_t k = 0; k < p.npoints; k++) {
                for (size_t v = 0; v < stride; v++) out += (v ? "," : "") + shortest_e(p.data[k * stride + v]);
                out += "\n";
            }
        } else {
            out += "Title: " + title + "\nDate: Thu Jan  1 00:00:00 1970\nPlotname: " + p.plotname + "\n";
            out += p.complex ? "Flags: complex\n" : "Flags: real\n";
            out += "No. Variables: " + std::to_string(p.names.size()) + "\nNo. Points: " + std::to_string(p.npoints) +
                   "\nVariables:\n";
            for (size_t i = 0; i < p.names.size(); i++)
                out += "\t" + std::to_string(i) + "\t" + p.names[i] + "\t" + var_type(p.names[i]) + "\n";
            if (format == 3) {
                out += "Binary:\n";
                out.append(reinterpret_cast<const char *>(p.data.data()), p.data.size() * sizeof(double));
            } else {
                out += "Values:\n";
                for (size_t k = 0; k < p.npoints; k++) {
                    out += std::to_string(k);
                    const double *row = &p.data[k * stride];
                    if (p.complex) {
                        for (size_t v = 0; v < stride / 2; v++)
                            out += "\t" + shortest_e(row[2 * v]) + "," + shortest_e(row[2 * v + 1]) + "\n";
                    } else {
                        for (size_t v = 0; v < stride; v++) out += "\t" + shortest_e(row[v]) + "\n";
                    }
                }
            }
        }
    }
    return out;
}

std::string problem_print_tables(espice_problem *problem, int format = 0)
{
    Problem *P = get(problem);
    if (!P) return "";
    return encode_plots(problem_plots(P), P->bd.deck.title, format);
}

// The measures of point 0 as JSON: {"name": value, ...} (null for NaN).
std::string problem_measures_json(espice_problem *problem)
{
    Problem *P = get(problem);
    if (!P || P->points.empty() || !P->points[0].ran) return "{}";
    const Point &pt = P->points[0];
    std::string out = "{";
    for (size_t k = 0; k < P->bd.deck.meas.size(); k++) {
        const double v = pt.failed ? NAN : pt.meas[k];
        char b[64];
        if (std::isnan(v)) std::snprintf(b, sizeof b, "null");
        else std::snprintf(b, sizeof b, "%.17g", v);
        out += std::string(k ? ", " : "") + "\"" + P->bd.deck.meas[k].name + "\": " + b;
    }
    return out + "}";
}

}  // namespace esp

extern "C" espice_status espice_cli_report(espice_problem *problem, int mode, char *buf, size_t cap, size_t *required)
{
    if (!get(problem)) return ESPICE_INVALID_HANDLE;
    const std::string s = mode == 1 ? esp::problem_measures_json(problem)
                        : mode >= 2 && mode <= 5 ? esp::problem_print_tables(problem, mode - 2)
                                                 : esp::problem_report(problem);
    if (required) *required = s.size() + 1;
    if (cap < s.size() + 1) return ESPICE_BUFFER_TOO_SMALL;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return ESPICE_OK;
}

extern "C" int espice_spice_number(const char *text, double *out)
{
    if (!text || !out) return 0;
    std::string t(text);
    for (char &c : t) c = (char)std::tolower((unsigned char)c);
    return esp::spice_number(t, *out) ? 1 : 0;
}

extern "C" espice_status espice_lanes_mc_draw(uint64_t seed, uint64_t site, uint32_t first_trial, uint32_t count,
                                              uint32_t fn, const double *args, uint32_t n_args, double *out)
{
    if (!args || n_args < 1 || (count && !out) || fn > esp::MC_LIMIT) return ESPICE_INVALID_ARGUMENT;
    for (uint32_t l = 0; l < count; l++) out[l] = esp::trial_draw(seed, first_trial + l, site, (int)fn, args, n_args);
    return ESPICE_OK;
}

extern "C" uint64_t espice_lanes_site_key(uint64_t owner, uint64_t salt) { return esp::site_key(owner, salt); }
