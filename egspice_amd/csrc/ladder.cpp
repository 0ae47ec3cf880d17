// The operating point's ladder over lanes (host driver): the plain rung on
// every lane, then the reference's stepping rungs for the lanes it leaves
// unconverged or singular (src/analysis/dc/op.zig:140-330 solveLadder and
// afterPlain, the host half of solveOnDevice :660-688). The reference runs
// each failed lane's ladder serially on the CPU; here each lane keeps its
// own ladder state on the host and every solve of every lane still
// climbing runs in one homotopy launch (espice_lanes_newton_homotopy), so
// a round of the ladder is one kernel over the lanes that need it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/espice_lanes.h"

namespace {

// the deck's gmin, the descent's target and the compiled kinds'
// $simparam("gmin") outside the rungs (src/core/numerics.zig:386)
constexpr double GMIN_DECK = 1e-12;
// solves per stepping rung at most (op.zig:333 max_ladder_solves)
constexpr uint32_t MAX_LADDER_SOLVES = 1000;
// a descent keeping its factor under 10^(1/20) for 20 solves in a row is
// stalled (op.zig:336-346); the double nearest 10^(1/20)
constexpr uint32_t STALL_SOLVES = 20;
constexpr double STALL_FACTOR = 1.1220184543019633;

// op.zig:383-387 gminFactor: an easy step (a quarter of the budget or
// less) accelerates, capped at 10; a hard one (over three quarters)
// slows, floored at 1.00005.
double gmin_factor(double factor, uint32_t iterations, uint32_t budget)
{
    if (iterations <= budget / 4) return std::min(factor * std::sqrt(factor), 10.0);
    if (iterations > 3 * (budget / 4)) return std::max(std::sqrt(factor), 1.00005);
    return factor;
}

struct Step {
    double gmin;
    bool converged;
};

// op.zig:352-366 gminStalled: the last STALL_SOLVES solves all kept the
// last good gmin over the next gmin tried under STALL_FACTOR.
bool gmin_stalled(const std::vector<Step> &steps, double start, double next)
{
    if (steps.size() < STALL_SOLVES) return false;
    double good = start;
    for (size_t i = 0; i < steps.size(); i++) {
        if (steps[i].converged) good = steps[i].gmin;
        if (i < steps.size() - STALL_SOLVES) continue;
        const double following = i + 1 < steps.size() ? steps[i + 1].gmin : next;
        if (good / following >= STALL_FACTOR) return false;
    }
    return true;
}

enum Phase { GMIN, CLEAN, SRC, SRC_OFF, FINAL, DONE };

struct LaneLadder {
    Phase ph = DONE;
    uint32_t total = 0;  // iterations of every solve, saturating at 65535
    uint32_t rung = 0;
    bool converged = false;
    // rung 2 (dynamic gmin stepping)
    double factor = 10.0, good_gmin = 0, gmin_val = 0, tried = 0;
    bool have_good = false;
    uint32_t solves = 0;
    std::vector<Step> steps;
    // rung 3 (source stepping) and its sources-off gmin rescue
    double lambda = 0, lambda_good = -1, delta = 0.25, off_g = 0;
    uint32_t src_solves = 0, off_i = 0;
    std::vector<double> x_good;

    void add(uint32_t it) { total = std::min<uint32_t>(total + it, 65535u); }
};

class Driver {
public:
    Driver(uint32_t n, uint32_t N, double gmin_start, uint32_t budget, uint32_t itl1)
        : n_(n), N_(N), gmin_start_(gmin_start), budget_(budget), itl1_(itl1), x_((size_t)n * N), lanes_(N)
    {
    }

    std::vector<double> &x() { return x_; }
    std::vector<LaneLadder> &lanes() { return lanes_; }

    // rung 2's start (op.zig:185-199): from a cold start, gmin_start / 10
    void start_gmin(uint32_t l)
    {
        LaneLadder &s = lanes_[l];
        cold(l);
        s.ph = GMIN;
        s.factor = 10.0;
        s.good_gmin = gmin_start_;
        s.gmin_val = s.good_gmin / s.factor;
        s.have_good = false;
        s.solves = 0;
        s.steps.clear();
        s.x_good.assign(n_, 0.0);
    }

    // The control of lane l's next solve and that solve's Newton budget.
    espice_lane_homotopy request(uint32_t l, uint32_t &max_iter)
    {
        LaneLadder &s = lanes_[l];
        espice_lane_homotopy h{0.0, GMIN_DECK, 1.0, ESPICE_HOM_ACTIVE, 0};
        max_iter = budget_;
        switch (s.ph) {
        case GMIN:
            s.tried = s.gmin_val;
            h.gshunt = s.gmin_val;
            h.gmodel = s.gmin_val;
            if (!s.have_good) h.flags |= ESPICE_HOM_COLD;
            break;
        case CLEAN:
            max_iter = itl1_;
            break;
        case SRC:
            h.srcfact = s.lambda;
            if (s.lambda_good < 0.0) h.flags |= ESPICE_HOM_COLD;
            break;
        case SRC_OFF:
            h.gshunt = s.off_g;
            h.srcfact = s.lambda;
            if (s.off_i == 0) h.flags |= ESPICE_HOM_COLD;
            break;
        case FINAL:
            max_iter = itl1_;
            break;
        case DONE:
            h.flags = 0;
            break;
        }
        return h;
    }

    // Lane l's solve ended with `r` (x_ holds where it left the lane).
    void respond(uint32_t l, const espice_lane_result &r)
    {
        LaneLadder &s = lanes_[l];
        const bool singular = (r.status & 2u) != 0;
        const bool conv = (r.status & 1u) != 0 && !singular;
        const uint32_t it = singular ? 0 : r.iterations;
        s.add(it);
        switch (s.ph) {
        case GMIN: gmin_result(l, conv, singular, it); break;
        case CLEAN:
            if (conv) finish(l, true, 2);
            else start_src(l);
            break;
        case SRC:
            if (!conv && s.lambda_good < 0.0) {
                // op.zig:393-415 sourcesOffGmin: the sources-off point again
                // under a diagonal gmin from 1e10 times the deck's down to
                // the deck's, a decade per solve, each from the last
                cold(l);
                s.off_g = GMIN_DECK * 1e10;
                s.off_i = 0;
                s.ph = SRC_OFF;
            } else {
                src_result(l, conv);
            }
            break;
        case SRC_OFF:
            if (!conv) {
                src_result(l, false);
            } else {
                s.off_g /= 10;
                if (++s.off_i == 11) src_result(l, true);
            }
            break;
        case FINAL: finish(l, conv, conv ? 3 : 0); break;
        case DONE: break;
        }
    }

private:
    void cold(uint32_t l)
    {
        for (uint32_t i = 0; i < n_; i++) x_[(size_t)i * N_ + l] = 0.0;
    }
    void save(uint32_t l)
    {
        LaneLadder &s = lanes_[l];
        for (uint32_t i = 0; i < n_; i++) s.x_good[i] = x_[(size_t)i * N_ + l];
    }
    void restore(uint32_t l)
    {
        const LaneLadder &s = lanes_[l];
        for (uint32_t i = 0; i < n_; i++) x_[(size_t)i * N_ + l] = s.x_good[i];
    }
    void finish(uint32_t l, bool conv, uint32_t rung)
    {
        LaneLadder &s = lanes_[l];
        s.converged = conv;
        s.rung = rung;
        s.ph = DONE;
    }

    // op.zig:212-252, one solve of the descent
    void gmin_result(uint32_t l, bool conv, bool refused, uint32_t it)
    {
        LaneLadder &s = lanes_[l];
        if (conv) {
            if (s.gmin_val <= GMIN_DECK) {
                // the last solve drops the shunt (ngspice's dynamic_gmin)
                s.ph = CLEAN;
                return;
            }
            save(l);
            s.have_good = true;
            s.good_gmin = s.gmin_val;
            s.factor = gmin_factor(s.factor, it, budget_);
            if (s.gmin_val < s.factor * GMIN_DECK) {
                s.factor = s.gmin_val / GMIN_DECK;
                s.gmin_val = GMIN_DECK;
            } else {
                s.gmin_val /= s.factor;
            }
        } else {
            if (s.factor < 1.00005) {
                start_src(l);  // wedged against the last good step
                return;
            }
            // a matrix the factor refused backs up by the square root only
            s.factor = refused ? std::sqrt(s.factor) : std::sqrt(std::sqrt(s.factor));
            s.gmin_val = s.good_gmin / s.factor;
            if (s.have_good) restore(l);
            else cold(l);
        }
        s.steps.push_back(Step{s.tried, conv});
        if (gmin_stalled(s.steps, gmin_start_, s.gmin_val) || ++s.solves >= MAX_LADDER_SOLVES) start_src(l);
    }

    // rung 3's start (op.zig:257-270): sources off, from a cold start
    void start_src(uint32_t l)
    {
        LaneLadder &s = lanes_[l];
        cold(l);
        s.ph = SRC;
        s.lambda = 0.0;
        s.lambda_good = -1.0;
        s.delta = 0.25;
        s.src_solves = 0;
    }

    // op.zig:291-307: grow the step 1.5x on success, halve it on failure
    // and retry from the last good lambda and x
    void src_result(uint32_t l, bool conv)
    {
        LaneLadder &s = lanes_[l];
        s.ph = SRC;
        if (conv) {
            if (s.lambda >= 1.0) {
                s.ph = FINAL;  // full sources reached
                return;
            }
            s.lambda_good = s.lambda;
            save(l);
            s.delta *= 1.5;
            s.lambda = std::min(s.lambda + s.delta, 1.0);
        } else {
            s.delta *= 0.5;
            if (s.delta < 1e-4 || s.lambda_good < 0.0) {
                s.ph = FINAL;
                return;
            }
            restore(l);
            s.lambda = std::min(s.lambda_good + s.delta, 1.0);
        }
        if (++s.src_solves >= MAX_LADDER_SOLVES) s.ph = FINAL;
    }

    uint32_t n_, N_;
    double gmin_start_;
    uint32_t budget_, itl1_;
    std::vector<double> x_;
    std::vector<LaneLadder> lanes_;
};


// Rounds of the ladder until every lane the driver started is done: each
// round uploads the host's x, runs every climbing lane's next solve (one
// launch per Newton budget the round needs: the stepping solves' and the
// plain one's) and hands each lane its outcome.
espice_status climb(espice_lanes *L, const espice_newton_opts *opts, const espice_ladder_opts *lo, Driver &d,
                    uint32_t N, double &ms_all)
{
    std::vector<espice_lane_homotopy> hom(N), part(N);
    std::vector<espice_lane_result> res(N), r(N);
    std::vector<uint32_t> budget(N);
    bool any = false;
    for (uint32_t l = 0; l < N; l++) any = any || d.lanes()[l].ph != DONE;
    while (any) {
        for (uint32_t l = 0; l < N; l++) hom[l] = d.request(l, budget[l]);
        if (espice_status e = espice_lanes_upload_x(L, d.x().data())) return e;
        const uint32_t budgets[2] = {lo->max_iter, opts->max_iter};
        const int n_budgets = lo->max_iter == opts->max_iter ? 1 : 2;
        for (int b = 0; b < n_budgets; b++) {
            bool some = false;
            for (uint32_t l = 0; l < N; l++) {
                part[l] = hom[l];
                if (n_budgets == 2 && budget[l] != budgets[b]) part[l].flags = 0;
                some = some || (part[l].flags & ESPICE_HOM_ACTIVE);
            }
            if (!some) continue;
            espice_newton_opts o = *opts;
            o.max_iter = budgets[b];
            o.gmin = 0.0;
            o.init_fix = 0;
            double ms = 0.0;
            if (espice_status e = espice_lanes_newton_homotopy(L, &o, part.data(), nullptr, r.data(), &ms)) return e;
            ms_all += ms;
            for (uint32_t l = 0; l < N; l++)
                if (part[l].flags & ESPICE_HOM_ACTIVE) res[l] = r[l];
        }
        // the lanes that ran leave their x on the device; the rest keep
        // the host's copy, which upload_x wrote there unchanged
        if (espice_status e = espice_lanes_download_x(L, d.x().data())) return e;
        any = false;
        for (uint32_t l = 0; l < N; l++) {
            if (!(hom[l].flags & ESPICE_HOM_ACTIVE)) continue;
            d.respond(l, res[l]);
            any = any || d.lanes()[l].ph != DONE;
        }
    }
    return ESPICE_OK;
}

// The plain rung's outcome of lane l: converged lanes are done (rung 1),
// the others start the gmin descent.
void after_plain(Driver &d, uint32_t l, const espice_lane_result &r)
{
    LaneLadder &s = d.lanes()[l];
    const bool singular = (r.status & 2u) != 0;
    s.total = singular ? 0 : std::min<uint32_t>(r.iterations, 65535u);
    s.steps.clear();
    if ((r.status & 1u) && !singular) {
        s.converged = true;
        s.rung = 1;
        s.ph = DONE;
        return;
    }
    s.converged = false;
    d.start_gmin(l);
}

}  // namespace

extern "C" espice_status espice_lanes_op_ladder(espice_lanes *L, const espice_newton_opts *opts,
                                                const espice_ladder_opts *lo, espice_lane_result *out,
                                                uint32_t *rung, double *kernel_ms)
{
    if (!L || !opts || !lo || !out) return ESPICE_INVALID_ARGUMENT;
    if (!(lo->gmin_start > 0.0) || lo->max_iter == 0 || opts->max_iter == 0) return ESPICE_INVALID_ARGUMENT;
    uint32_t n = 0, N = 0;
    if (espice_status e = espice_lanes_shape(L, &n, &N)) return e;
    double ms_all = 0.0, ms = 0.0;
    // rung 1: the plain Newton on every lane, as the caller's options say
    if (espice_status e = espice_lanes_newton_op(L, opts, out, &ms)) return e;
    ms_all += ms;
    Driver d(n, N, lo->gmin_start, lo->max_iter, opts->max_iter);
    if (espice_status e = espice_lanes_download_x(L, d.x().data())) return e;
    for (uint32_t l = 0; l < N; l++) after_plain(d, l, out[l]);
    if (espice_status e = climb(L, opts, lo, d, N, ms_all)) return e;
    for (uint32_t l = 0; l < N; l++) {
        const LaneLadder &s = d.lanes()[l];
        out[l] = espice_lane_result{s.converged ? 1u : 0u, s.total};
        if (rung) rung[l] = s.converged ? s.rung : 0;
    }
    if (kernel_ms) *kernel_ms = ms_all;
    return ESPICE_OK;
}

extern "C" espice_status espice_lanes_dc_sweep(espice_lanes *L, uint32_t kind, uint32_t instance,
                                               const double *values, uint32_t n_points,
                                               const espice_newton_opts *opts, const espice_ladder_opts *lo,
                                               double *x_out, uint32_t *converged, double *kernel_ms)
{
    if (!L || !values || !opts || !lo || !x_out || !converged) return ESPICE_INVALID_ARGUMENT;
    if (!(lo->gmin_start > 0.0) || lo->max_iter == 0 || opts->max_iter == 0) return ESPICE_INVALID_ARGUMENT;
    uint32_t n = 0, N = 0;
    if (espice_status e = espice_lanes_shape(L, &n, &N)) return e;
    const size_t nx = (size_t)n * N;
    double ms_all = 0.0;
    Driver d(n, N, lo->gmin_start, lo->max_iter, opts->max_iter);
    // sol (d.x) and the previous point per lane, both node-major
    std::vector<double> &x = d.x();
    std::vector<double> prev(nx, 0.0);
    std::vector<uint8_t> cold(N, 1);
    std::vector<espice_lane_homotopy> hom(N);
    std::vector<espice_lane_result> r(N);
    for (uint32_t pt = 0; pt < n_points; pt++) {
        if (espice_status e = espice_lanes_set_source(L, kind, instance, values[pt])) return e;
        // MODEINITPRED (dc.zig:463-475): the last two points extrapolated,
        // the rotation running on cold points too
        for (size_t k = 0; k < nx; k++) {
            const double last = x[k];
            x[k] = 2.0 * last - prev[k];
            prev[k] = last;
        }
        // the warm lanes' Newton from the predictor limited against the last
        // point (prepareWarm), with the `.dc` budget (warmOptions: itl2)
        bool any_warm = false;
        for (uint32_t l = 0; l < N; l++) {
            hom[l] = espice_lane_homotopy{0.0, GMIN_DECK, 1.0, 0u, 0u};
            if (!cold[l]) {
                hom[l].flags = ESPICE_HOM_ACTIVE | ESPICE_HOM_PREDICT;
                any_warm = true;
            }
        }
        std::vector<uint8_t> done(N, 0);
        if (any_warm) {
            if (espice_status e = espice_lanes_upload_x(L, x.data())) return e;
            espice_newton_opts o = *opts;
            o.max_iter = lo->max_iter;
            o.gmin = 0.0;
            o.init_fix = 0;
            double ms = 0.0;
            if (espice_status e = espice_lanes_newton_homotopy(L, &o, hom.data(), prev.data(), r.data(), &ms))
                return e;
            ms_all += ms;
            if (espice_status e = espice_lanes_download_x(L, x.data())) return e;
            for (uint32_t l = 0; l < N; l++)
                if (!cold[l] && (r[l].status & 1u) && !(r[l].status & 2u)) done[l] = 1;
        }
        // every other lane: the cold ladder from zero (endPoint's
        // coldStart + nodesetLadder), its plain rung one homotopy launch
        bool any_cold = false;
        for (uint32_t l = 0; l < N; l++) {
            hom[l] = espice_lane_homotopy{0.0, GMIN_DECK, 1.0, 0u, 0u};
            d.lanes()[l].ph = DONE;
            if (done[l]) continue;
            for (uint32_t i = 0; i < n; i++) x[(size_t)i * N + l] = 0.0;
            hom[l].flags = ESPICE_HOM_ACTIVE | ESPICE_HOM_COLD;
            any_cold = true;
        }
        if (any_cold) {
            if (espice_status e = espice_lanes_upload_x(L, x.data())) return e;
            espice_newton_opts o = *opts;
            o.gmin = 0.0;
            o.init_fix = 0;
            double ms = 0.0;
            if (espice_status e = espice_lanes_newton_homotopy(L, &o, hom.data(), nullptr, r.data(), &ms)) return e;
            ms_all += ms;
            if (espice_status e = espice_lanes_download_x(L, x.data())) return e;
            for (uint32_t l = 0; l < N; l++)
                if (!done[l]) after_plain(d, l, r[l]);
            if (espice_status e = climb(L, opts, lo, d, N, ms_all)) return e;
            for (uint32_t l = 0; l < N; l++)
                if (!done[l] && d.lanes()[l].converged) done[l] = 1;
        }
        // endPoint: a converged point is recorded and the next one warm;
        // the first point's solution is also the predictor's previous one
        // (dctrcurv.c:457-462, dc.zig:524-527)
        for (uint32_t l = 0; l < N; l++) {
            converged[(size_t)pt * N + l] = done[l];
            for (uint32_t i = 0; i < n; i++)
                x_out[((size_t)pt * n + i) * N + l] = done[l] ? x[(size_t)i * N + l] : NAN;
            cold[l] = done[l] ? 0 : 1;
            if (done[l] && pt == 0)
                for (uint32_t i = 0; i < n; i++) prev[(size_t)i * N + l] = x[(size_t)i * N + l];
        }
    }
    if (kernel_ms) *kernel_ms = ms_all;
    return ESPICE_OK;
}
