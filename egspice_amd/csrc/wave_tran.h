// One lane's transient on its wavefront, in one launch: the operating
// point, then the step machine of src/solver/kernels/lane.zig (stepBegin,
// stepEndA, stepEndB, which restate tran.Sim.begin/end, tran.zig:749-1087,
// for order <= 2), the integrator's companion and LTE bound
// (integrator.zig companionAt, stepBound), the sources' breakpoints
// (device/eval.zig spicePulseBreak behind srcBreak's caches) and the
// .meas passes of src/output/scan.zig fed with every accepted row.
//
// Thread 0 runs the scalar control (accept, reject, order, dt, landing,
// the passes) and publishes its decisions in LDS; the whole wavefront runs
// the Newton, the predictor, the charge evaluation, advanceCurrent and the
// row copies. The reference runs the same machine as ~10 kernels per round
// with the host draining accepted rows; here nothing leaves the CU until
// the lane is done.
#pragma once

#include "wave_newton.h"

namespace esp {

struct TranOpts {
    double t_stop, dt_init, dt_max, dt_min, xmu;
    double reltol, abstol, vntol, chgtol, trtol;
    uint32_t method, max_iter, max_steps, record_cap;
    // op_warm: the operating point starts at the lane's x; mixed_refine
    // (the host's switch): FP32 linear steps with at most this many solves
    uint32_t op_warm, mixed_refine;
    // early_stop: the lane ends once every pass has stopped; abort_after
    // (NaN off): past it, a pass of abort_mask still running aborts the lane
    uint32_t early_stop, pad2;
    double abort_after;
    uint64_t abort_mask;
    double op_reltol;  // the operating point's reltol when positive (host side)
};

// One .meas pass (scan.zig): 0 Cross (WHEN/TRIG/TARG), 1 At (FIND, or
// DERIV with op 1), 2 Extremum (op 0 MIN, 1 MAX, 2 AVG), 3 Integ (op 0
// INTEG, 1 RMS), 4 Find (FIND, or DERIV with op 1, ... WHEN: the pass before
// it is its WHEN's Cross). `row` is the unknown it reads.
struct MeasPass {
    uint32_t kind, row;
    int32_t rise, fall, cross;
    uint32_t op;
    double val, td, from, to, at;
};

struct MeasOut {
    double value, aux0, aux1;  // Cross: time; At: value; Extremum: value, at; Integ: value, from, to
};

struct TranOut {
    uint32_t status;  // 0 reached t_stop (or stopped early), 1 dt underflow, 2 no operating point, 3 singular
                      // operating point, 4 aborted by the measure bound
    uint32_t points, steps, attempts, rej_newton, rej_lte, bp_landings, nr_iters;
    // the mixed-precision linear steps (zero in FP64): FP32 solves taken,
    // Newton ticks handed back to the FP64 factor
    uint32_t fp32_solves, fp64_fallbacks;
};

constexpr uint32_t MEAS_WORDS = 24;  // LDS words of one pass's state
constexpr uint32_t TRAN_SCALARS = 32;

// Extra 8-byte words of a transient lane: see tran_words in wave.h, plus
// two breakpoint-cache words per instance and one state block per pass.
ESP_HD size_t tran_extra_words(uint32_t n_inst, uint32_t n_pass) { return 2 * (size_t)n_inst + MEAS_WORDS * (size_t)n_pass; }

// -------------------------------------------------------------- helpers

ESP_HD bool ulps_close(double a, double b, int64_t max_ulps)
{
    if (a == b) return true;
    const int64_t ia = (int64_t)bits(a), ib = (int64_t)bits(b);
    const int64_t la = ia < 0 ? INT64_MIN - ia : ia;
    const int64_t lb = ib < 0 ? INT64_MIN - ib : ib;
    int64_t d;
    if (__builtin_sub_overflow(la, lb, &d)) return false;
    return (d < 0 ? -d : d) <= max_ulps;
}

ESP_HD double sel_max(double a, double b) { return a > b ? a : b; }
ESP_HD double zmin(double a, double b) { return a < b ? a : b; }  // @min on non-NaN operands

struct TranCoeffs {
    int method;
    double ag0, ag1, ag2;
};

// integrator.zig coeffs.
ESP_HD TranCoeffs tran_coeffs(int method, double dt, double dt_prev, double xmu)
{
    TranCoeffs c{method, 0.0, 1.0, 0.0};
    if (method == 0) {
        c.ag0 = 1.0 / dt;
    } else if (method == 1) {
        c.ag0 = 1.0 / dt / (1.0 - xmu);
        c.ag1 = xmu / (1.0 - xmu);
    } else {
        const double r = dt / dt_prev;
        c.ag0 = (1.0 + 2.0 * r) / ((1.0 + r) * dt);
        c.ag2 = r * r / ((1.0 + r) * dt);
    }
    return c;
}

struct LteIn {
    double dt, dt1, dt2, reltol, abstol, chgtol, trtol;
};

// integrator.zig stepBound / minDel: one charge site's bound; `bad` is set
// when one of its inputs or the bound is NaN (the reference poisons the
// whole bound then).
ESP_HD double step_site(int method, int cur_method, const double *q0, const double *q1, const double *q2,
                        const double *q3, const double *ip, uint32_t j, const TranCoeffs &c, const LteIn &e,
                        bool &bad)
{
    const double coeff = method == 0 ? 0.5 : method == 1 ? 0.08333333333 : 0.2222222222;
    const double sum01 = e.dt + e.dt1, sum12 = e.dt1 + e.dt2, sum012 = e.dt + (e.dt1 + e.dt2);
    const double qc = q0[j], qp = q1[j], qp2 = q2[j], i_p = ip[j];
    double i_new;
    if (cur_method == 1) i_new = c.ag0 * (qc - qp) - c.ag1 * i_p;
    else if (cur_method == 2) i_new = c.ag0 * (qc - qp) - c.ag2 * (qp - qp2);
    else i_new = c.ag0 * (qc - qp);
    const double volttol = e.abstol + e.reltol * sel_max(fabs_(i_new), fabs_(i_p));
    const double chargetol = e.reltol * sel_max(sel_max(fabs_(qc), fabs_(qp)), e.chgtol) / e.dt;
    const double tol = sel_max(volttol, chargetol);
    // (x - x) is 0, or NaN for a NaN or infinite x
    double poison = (qc - qc) + (qp - qp) + (qp2 - qp2) + (i_p - i_p);
    const double f12 = (qp - qp2) / e.dt1;
    double dd = ((qc - qp) / e.dt - f12) / sum01;
    if (method != 0) {
        const double qp3 = q3[j];
        poison += qp3 - qp3;
        const double f123 = (f12 - (qp2 - qp3) / e.dt2) / sum12;
        dd = (dd - f123) / sum012;
    }
    const double del = e.trtol * tol / sel_max(e.abstol, coeff * fabs_(dd));
    bad = bad || poison != poison || del != del;
    return del;
}

// The bound over every site: their minimum, NaN when any site is poisoned,
// the square root of it for the second-order methods.
ESP_HD double step_bound(int method, int cur_method, const double *q0, const double *q1, const double *q2,
                         const double *q3, const double *ip, uint32_t len, const TranCoeffs &c, const LteIn &e)
{
    double vmin = inf_();
    bool bad = false;
    for (uint32_t j = 0; j < len; j++) {
        const double del = step_site(method, cur_method, q0, q1, q2, q3, ip, j, c, e, bad);
        vmin = del < vmin ? del : vmin;
    }
    if (bad) return __builtin_nan("");
    return method == 0 ? vmin : __builtin_sqrt(vmin);
}

#if defined(__HIPCC__)
// step_bound across the wavefront (all threads, one wave): thread t takes
// sites t, t + 64, ...; a minimum and a flag do not depend on the order the
// sites are visited in, so every thread returns step_bound's bits.
__device__ inline double wave_step_bound(int method, int cur_method, const double *q0, const double *q1,
                                         const double *q2, const double *q3, const double *ip, uint32_t len,
                                         const TranCoeffs &c, const LteIn &e)
{
    double vmin = inf_();
    bool bad = false;
    for (uint32_t j = threadIdx.x; j < len; j += blockDim.x) {
        const double del = step_site(method, cur_method, q0, q1, q2, q3, ip, j, c, e, bad);
        vmin = del < vmin ? del : vmin;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(vmin, off);
        vmin = o < vmin ? o : vmin;
    }
    if (__ballot(bad ? 1 : 0) != 0) return __builtin_nan("");
    return method == 0 ? vmin : __builtin_sqrt(vmin);
}
#endif

// --------------------------------------------------------- .meas passes

// scan.zig Cross/At/Extremum/Integ state, in one LDS block per pass.
struct MeasState {
    double d[16];
    int32_t i[8];
};
static_assert(sizeof(MeasState) <= 8 * MEAS_WORDS, "pass state fits its LDS block");

ESP_HD void meas_init(const MeasPass &m, MeasState &s)
{
    for (int k = 0; k < 16; k++) s.d[k] = 0.0;
    for (int k = 0; k < 8; k++) s.i[k] = 0;
    if (m.kind == 0) {
        s.d[2] = __builtin_nan("");  // last
        s.d[3] = __builtin_nan("");  // the event
    } else if (m.kind == 1 || m.kind == 4) {
        s.d[2] = __builtin_nan("");  // value
    } else if (m.kind == 3) {
        s.d[4] = __builtin_nan("");  // first_x
        s.d[5] = __builtin_nan("");  // cut_x
    }
}

// Integ.panel (scan.zig:311-331) over the pending points d[8..15]: Simpson
// 3/8 on three equal widths, Simpson 1/3 on two, else the trapezoid, into
// the sums d[1..3].
ESP_HD void integ_panel(MeasState &s, uint32_t avail)
{
    double p[4][2];
    for (uint32_t k = 0; k < 4; k++) {
        p[k][0] = k < avail ? s.d[8 + 2 * k] : 0.0;
        p[k][1] = k < avail ? s.d[9 + 2 * k] : 0.0;
    }
    double width[3];
    for (uint32_t j = 0; j < 3; j++) width[j] = j + 1 < avail ? p[j + 1][0] - p[j][0] : 0;
    uint32_t used;
    if (ulps_close(width[0], width[1], 100) && ulps_close(width[0], width[2], 100)) {
        s.d[1] += 3 * width[0] * (p[0][1] + 3 * (p[1][1] + p[2][1]) + p[3][1]) / 8.0;
        used = 3;
    } else if (ulps_close(width[0], width[1], 100)) {
        s.d[2] += width[0] * (p[0][1] + 4 * p[1][1] + p[2][1]) / 3.0;
        used = 2;
    } else {
        s.d[3] += width[0] * (p[0][1] + p[1][1]) / 2;
        used = 1;
    }
    const uint32_t keep = avail > used ? avail - used : 0;
    for (uint32_t k = 0; k < keep; k++) {
        s.d[8 + 2 * k] = s.d[8 + 2 * (avail - keep + k)];
        s.d[9 + 2 * k] = s.d[9 + 2 * (avail - keep + k)];
    }
    s.i[2] = (int32_t)keep;
}

ESP_HD double meas_line(double x0, double y0, double x1, double y1, double at)
{
    const double slope = (y1 - y0) / (x1 - x0);
    const double intercept = y0 - slope * x0;
    return slope * at + intercept;
}

// Feeds sample (xv, v); s.i[7] set once the pass has stopped.
ESP_HD void meas_feed(const MeasPass &m, MeasState &s, double xv, double v)
{
    if (s.i[7]) return;
    if (m.kind == 0) {
        // Cross: d0 prev, d1 prev_x, d2 last, d3 event; i0 rise, i1 fall,
        // i2 cross, i3 above, i4 pending, i5 first, i6 event found
        if (xv < m.td || xv < m.from) return;
        if (m.to != 0 && xv > m.to) {
            s.i[7] = 1;
            return;
        }
        const double level = m.val, prev_level = m.val;
        if (s.i[5] == 1) {
            s.i[2] = 0;
            s.i[3] = !(v < level);
            if (!s.i[3] && s.d[0] >= prev_level) {
                s.i[1] = 1;
                s.i[2] = 1;
            } else if (s.i[3] && s.d[0] < prev_level) {
                s.i[0] = 1;
                s.i[2] = 1;
            }
        }
        if (s.i[5] > 1) {
            if (!s.i[3] && v >= level) {
                s.i[3] = 1;
                s.i[2] += 1;
                s.i[0] += 1;
                if (m.fall != -2) s.i[4] = 1;
            } else if (s.i[3] && v <= level) {
                s.i[3] = 0;
                s.i[2] += 1;
                s.i[1] += 1;
                if (m.rise != -2) s.i[4] = 1;
            }
            const double hit = s.d[1] + (prev_level - s.d[0]) * (xv - s.d[1]) / (v - s.d[0] - level + prev_level);
            if (s.i[2] == m.cross || s.i[0] == m.rise || s.i[1] == m.fall) {
                s.d[3] = hit;
                s.i[6] = 1;
                s.i[7] = 1;
                return;
            }
            if (s.i[4]) {
                if (m.cross == -1 && m.rise == -1 && m.fall == -1) {
                    s.d[3] = hit;
                    s.i[6] = 1;
                    s.i[7] = 1;
                    return;
                }
                if (m.cross == -2 || m.rise == -2 || m.fall == -2) s.d[2] = hit;
                s.i[4] = 0;
            }
        }
        s.i[5] += 1;
        s.d[0] = v;
        s.d[1] = xv;
    } else if (m.kind == 1) {
        // At: d0 px, d1 pv, d2 value; i0 rows
        if (s.i[0] > 0 && s.d[0] <= m.at && xv >= m.at) {
            s.d[2] = m.op ? (v - s.d[1]) / (xv - s.d[0]) : s.d[1] + (m.at - s.d[0]) * (v - s.d[1]) / (xv - s.d[0]);
            s.i[7] = 1;
            return;
        }
        s.i[0] += 1;
        s.d[0] = xv;
        s.d[1] = v;
        if (xv > m.at) s.i[7] = 1;
    } else if (m.kind == 2) {
        // Extremum: d0 m, d1 m_at, d2 pv, d3 px, d4 span, d5 last x; i0 first
        s.d[5] = xv;
        if (xv < m.from) return;
        if (m.to != 0 && xv > m.to) {
            s.i[7] = 1;
            return;
        }
        if (!s.i[0]) {
            s.i[0] = 1;
            s.d[0] = m.op == 2 ? 0 : v;
            s.d[1] = xv;
            s.d[2] = v;
            s.d[3] = xv;
            return;
        }
        if (m.op == 0) {
            if (v <= s.d[0]) {
                s.d[0] = v;
                s.d[1] = xv;
            }
        } else if (m.op == 1) {
            if (v >= s.d[0]) {
                s.d[0] = v;
                s.d[1] = xv;
            }
        } else {
            s.d[0] += 0.5 * (v + s.d[2]) * (xv - s.d[3]);
            s.d[4] += xv - s.d[3];
            s.d[2] = v;
            s.d[3] = xv;
        }
    } else {
        // Integ: sums d1..d3, d4 first_x, d5 cut_x, d6 px, d7 py, p in
        // d8..d15; i0 inside, i1 cut, i2 np, i3 rows, i4 count
        bool stop = false;
        if (!(!s.i[0] && xv < m.from)) {
            s.i[0] = 1;
            s.i[1] = m.to != 0 && xv > m.to;
            double pt0 = xv, pt1 = v;
            if (s.i[1]) {
                if (s.i[3] > 0 && !ulps_close(xv, m.to, 100)) {
                    pt0 = m.to;
                    pt1 = meas_line(s.d[6], s.d[7], xv, v, m.to);
                }
            } else if (s.i[4] == 0 && m.from != 0 && s.i[3] > 0 && !ulps_close(xv, m.from, 100)) {
                pt0 = m.from;
                pt1 = meas_line(s.d[6], s.d[7], xv, v, m.from);
            }
            if (m.op) pt1 = pt1 * pt1;
            if (s.i[4] == 0) s.d[4] = pt0;
            if (s.i[1]) s.d[5] = pt0;
            s.i[4] += 1;
            s.d[8 + 2 * s.i[2]] = pt0;
            s.d[9 + 2 * s.i[2]] = pt1;
            s.i[2] += 1;
            if (s.i[2] == 4) integ_panel(s, 4);
            stop = s.i[1] != 0;
        }
        s.i[3] += 1;
        s.d[6] = xv;
        s.d[7] = v;
        if (stop) s.i[7] = 1;
    }
}

// Find (scan.zig Find, :339-380) after its WHEN's Cross has read the same
// sample: until the WHEN ends it keeps the FIND vector's last three rows;
// at the WHEN's event it becomes an At there, fed those rows first, then
// the rows after. d0 px, d1 pv, d2 value, d4 the event, d8..d13 the rows
// (x, v) oldest first; i0 the At's rows, i1 rows kept, i2 state (0
// waiting, 1 armed, 2 no event), i3 rows fed.
ESP_HD bool find_at_feed(const MeasPass &m, MeasState &s, double xv, double v)
{
    if (s.i[0] > 0 && s.d[0] <= s.d[4] && xv >= s.d[4]) {
        s.d[2] = m.op ? (v - s.d[1]) / (xv - s.d[0]) : s.d[1] + (s.d[4] - s.d[0]) * (v - s.d[1]) / (xv - s.d[0]);
        return true;
    }
    s.i[0] += 1;
    s.d[0] = xv;
    s.d[1] = v;
    return false;
}

ESP_HD void meas_feed_find(const MeasPass &m, MeasState &s, const MeasState &when, double xv, double v)
{
    if (s.i[7]) return;
    if (s.i[2] == 1) {
        if (find_at_feed(m, s, xv, v) || xv > s.d[4]) s.i[7] = 1;
        return;
    }
    if (s.i[1] == 3)
        for (int k = 0; k < 4; k++) s.d[8 + k] = s.d[10 + k];
    s.i[1] = s.i[1] < 3 ? s.i[1] + 1 : 3;
    s.d[8 + 2 * (s.i[1] - 1)] = xv;
    s.d[9 + 2 * (s.i[1] - 1)] = v;
    s.i[3] += 1;
    if (!when.i[7]) return;
    const double time = when.i[6] ? when.d[3] : when.d[2];
    if (time != time) {
        s.i[2] = 2;
        s.i[7] = 1;
        return;
    }
    s.i[2] = 1;
    s.d[4] = time;
    if (s.i[3] > s.i[1] && time <= s.d[8]) {
        s.i[7] = 1;
        return;
    }
    for (int k = 0; k < s.i[1]; k++) {
        const double rx = s.d[8 + 2 * k], rv = s.d[9 + 2 * k];
        if (find_at_feed(m, s, rx, rv) || rx > time) {
            s.i[7] = 1;
            return;
        }
    }
}

ESP_HD MeasOut meas_result(const MeasPass &m, MeasState &s)
{
    MeasOut r{__builtin_nan(""), __builtin_nan(""), __builtin_nan("")};
    if (m.kind == 0) {
        r.value = s.i[6] ? s.d[3] : s.d[2];
    } else if (m.kind == 1 || m.kind == 4) {
        r.value = s.d[2];
    } else if (m.kind == 2) {
        if (m.op == 2) {
            r.value = s.d[0] / (s.i[0] ? s.d[4] : 1);
            r.aux0 = s.d[5];
        } else {
            r.value = s.d[0];
            r.aux0 = s.d[1];
        }
    } else {
        if (s.i[3] == 0) return r;
        while (s.i[2] > 1) integ_panel(s, (uint32_t)s.i[2]);
        const double from = s.i[4] > (s.i[1] ? 1 : 0) ? s.d[4] : __builtin_nan("");
        const double to = s.i[1] ? s.d[5] : s.d[6];
        const double total = s.d[1] + s.d[2] + s.d[3];
        r.value = m.op ? __builtin_sqrt(total / (to - from)) : total;
        r.aux0 = from;
        r.aux1 = to;
    }
    return r;
}

// ------------------------------------------------------- step machine

// One lane's step control (tran.zig Sim's scalars), run by thread 0.
struct TranStep {
    double t, dt, dt_old[6], dt_next, bp_next, bp_save_dt, bp_lo, bp_hi;
    double delmin, min_break, dt_att;
    uint32_t steps, order, run_order, order_run, perm[4];
    bool bp_clamped, stopped;
    TranOut out;
};

// The V sources' next breakpoint strictly after tq, asked at the accepted
// time t_acc: nextBreakpointFn's batch cache (device/eval.zig:2709-2730)
// over srcBreak's per-model caches (:1592-1598). bpc holds (req, at) per
// instance.
ESP_HD double tran_next_bp(const WaveTemplate &W, const double *prm, double *bpc, TranStep &st, double t_acc, double tq)
{
    if (st.bp_lo <= tq && tq < st.bp_hi) return st.bp_hi;
    double best = inf_();
    for (uint32_t e = 0; e < W.n_inst; e++) {
        const InstEntry ie = W.inst[e];
        if (ie.kind != K_VSRC && ie.kind != K_ISRC) continue;
        double *c = bpc + 2 * e;
        if (tq < c[0] || tq >= c[1]) {
            c[0] = tq;
            c[1] = src_break(LP{prm + ie.prm}, t_acc, tq);
        }
        best = c[1] < best ? c[1] : best;
    }
    st.bp_lo = tq;
    st.bp_hi = best;
    return best;
}

// floorRetry (tran.zig:1089-1105) after a rejection proposed st.dt.
ESP_HD void tran_floor_retry(const TranOpts &o, TranStep &st)
{
    if (st.dt < o.dt_min && st.dt_att > o.dt_min && st.t + o.dt_min > st.t) st.dt = o.dt_min;
    if (st.dt < o.dt_min) {
        st.out.status = 1;
        st.stopped = true;
    }
}

}  // namespace esp
