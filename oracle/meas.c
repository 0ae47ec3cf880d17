/* The .meas passes of a transient, over one recorded waveform. TEST
 * INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates src/output/scan.zig (ngspice com_measure2.c semantics), the
 * passes the reference runs on the host recording, on the streamed
 * transient and on its GPU lanes alike:
 *   or_meas_cross     Cross.feed :70-162 and finish :166-170 (WHEN, TRIG,
 *                     TARG: RISE/FALL/CROSS counts, TD, FROM/TO, LAST)
 *   or_meas_at        At.feed :174-195 (FIND/DERIV ... AT=)
 *   or_meas_extremum  Extremum.feed/stat :199-252 (MIN, MAX, AVG)
 *   or_meas_integ     Integ.feed/panel/result :261-346 (INTEG, RMS)
 * composed as src/output/measure.zig evaluate :948-1051 composes them.
 * Parity pinned by: tests/test_oracle_tran.py (docs/learn/measure.md:24-36,
 * the reference's printed values of docs/examples/measure.sp).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define UNSET (-1)
#define LAST (-2)

/* scan.zig almostEqualUlps :49-59 (100 ulps). */
static int ulps100(double a, double b)
{
    if (a == b) return 1;
    int64_t ia, ib;
    memcpy(&ia, &a, 8);
    memcpy(&ib, &b, 8);
    uint64_t oa = ia < 0 ? (uint64_t)INT64_MIN - (uint64_t)ia : (uint64_t)ia;
    uint64_t ob = ib < 0 ? (uint64_t)INT64_MIN - (uint64_t)ib : (uint64_t)ib;
    int64_t d = (int64_t)(oa - ob);
    if (d < 0) d = -d;
    return d <= 100;
}

/* scan.zig line :41-45: the line through two points at `at`. */
static double line(double x0, double y0, double x1, double y1, double at)
{
    const double slope = (y1 - y0) / (x1 - x0);
    const double intercept = y0 - slope * x0;
    return slope * at + intercept;
}

double or_meas_cross(const double *x, const double *y, uint32_t len, const or_meas_clause *cl)
{
    int32_t want_rise = cl->rise, want_fall = cl->fall, want_cross = cl->cross;
    int32_t rise = 0, fall = 0, cross = 0;
    int above = 0, pending = 0;
    double last = NAN, prev = 0, prev_x = 0, hit_out = NAN;
    uint32_t first = 0;
    int ended = 0, found = 0;
    const double level = cl->val, prev_level = cl->val;
    for (uint32_t i = 0; i < len; i++) {
        const double xv = x[i], v = y[i];
        if (xv < cl->td) continue;
        if (xv < cl->from) continue;
        if (cl->to != 0 && xv > cl->to) break;
        if (first == 1) {
            cross = 0;
            above = !(v < level);
            if (!above && prev >= prev_level) {
                fall = 1;
                cross = 1;
            } else if (above && prev < prev_level) {
                rise = 1;
                cross = 1;
            }
        }
        if (first > 1) {
            if (!above && v >= level) {
                above = 1;
                cross += 1;
                rise += 1;
                if (cl->fall != LAST) pending = 1;
            } else if (above && v <= level) {
                above = 0;
                cross += 1;
                fall += 1;
                if (cl->rise != LAST) pending = 1;
            }
            const double hit = prev_x + (prev_level - prev) * (xv - prev_x) / (v - prev - level + prev_level);
            if (cross == want_cross || rise == want_rise || fall == want_fall) {
                hit_out = hit;
                found = 1;
                ended = 1;
                break;
            }
            if (pending) {
                if (cl->cross == UNSET && cl->rise == UNSET && cl->fall == UNSET) {
                    hit_out = hit;
                    found = 1;
                    ended = 1;
                    break;
                }
                if (cl->cross == LAST || cl->rise == LAST || cl->fall == LAST) last = hit;
                pending = 0;
            }
        }
        first += 1;
        prev = v;
        prev_x = xv;
    }
    if (found) return hit_out;
    if (ended || isnan(last)) return NAN;
    return last;
}

double or_meas_at(const double *x, const double *y, uint32_t len, double at, int slope)
{
    double px = 0, pv = 0;
    for (uint32_t i = 0; i < len; i++) {
        const double xv = x[i], v = y[i];
        if (i > 0 && px <= at && xv >= at) return slope ? (v - pv) / (xv - px) : pv + (at - px) * (v - pv) / (xv - px);
        px = xv;
        pv = v;
    }
    return NAN;
}

void or_meas_extremum(const double *x, const double *y, uint32_t len, double from, double to, int op, double *value,
                      double *at)
{
    int first = 0;
    double m = 0, m_at = 0, pv = 0, px = 0, span = 0, xlast = 0;
    for (uint32_t i = 0; i < len; i++) {
        const double xv = x[i], v = y[i];
        xlast = xv;
        if (xv < from) continue;
        if (to != 0 && xv > to) break;
        if (!first) {
            first = 1;
            m = op == 2 ? 0 : v;
            m_at = xv;
            pv = v;
            px = xv;
            continue;
        }
        if (op == 0) {
            if (v <= m) {
                m = v;
                m_at = xv;
            }
        } else if (op == 1) {
            if (v >= m) {
                m = v;
                m_at = xv;
            }
        } else {
            m += 0.5 * (v + pv) * (xv - px);
            span += xv - px;
            pv = v;
            px = xv;
        }
    }
    if (op == 2) {
        *value = m / (first ? span : 1);
        *at = xlast;
    } else {
        *value = m;
        *at = m_at;
    }
}

typedef struct {
    double p[4][2];
    uint32_t np;
    double sum[3];
} integ_acc;

/* Integ.panel :311-331: Simpson 3/8 on three equal widths, Simpson 1/3 on
 * two, else the trapezoid; points past those available read zero. */
static void panel(integ_acc *s, uint32_t avail)
{
    double p[4][2];
    memcpy(p, s->p, sizeof p);
    for (uint32_t k = avail; k < 4; k++) p[k][0] = p[k][1] = 0;
    double width[3];
    for (uint32_t j = 0; j < 3; j++) width[j] = j + 1 < avail ? p[j + 1][0] - p[j][0] : 0;
    uint32_t used;
    if (ulps100(width[0], width[1]) && ulps100(width[0], width[2])) {
        s->sum[0] += 3 * width[0] * (p[0][1] + 3 * (p[1][1] + p[2][1]) + p[3][1]) / 8.0;
        used = 3;
    } else if (ulps100(width[0], width[1])) {
        s->sum[1] += width[0] * (p[0][1] + 4 * p[1][1] + p[2][1]) / 3.0;
        used = 2;
    } else {
        s->sum[2] += width[0] * (p[0][1] + p[1][1]) / 2;
        used = 1;
    }
    s->np = avail > used ? avail - used : 0;
    memmove(s->p[0], s->p[avail - s->np], sizeof(double[2]) * s->np);
}

void or_meas_integ(const double *x, const double *y, uint32_t len, double from, double to, int rms, double *value,
                   double *rfrom, double *rto)
{
    integ_acc s;
    memset(&s, 0, sizeof s);
    int inside = 0, cut = 0;
    uint64_t rows = 0, count = 0;
    double px = 0, py = 0, first_x = NAN, cut_x = NAN;
    for (uint32_t i = 0; i < len; i++) {
        const double xv = x[i], v = y[i];
        int stop = 0;
        if (!inside && xv < from) goto next;
        inside = 1;
        cut = to != 0 && xv > to;
        {
            double pt[2] = {xv, v};
            if (cut) {
                if (rows > 0 && !ulps100(xv, to)) {
                    pt[0] = to;
                    pt[1] = line(px, py, xv, v, to);
                }
            } else if (count == 0 && from != 0 && rows > 0 && !ulps100(xv, from)) {
                pt[0] = from;
                pt[1] = line(px, py, xv, v, from);
            }
            if (rms) pt[1] = pt[1] * pt[1];
            if (count == 0) first_x = pt[0];
            if (cut) cut_x = pt[0];
            count += 1;
            s.p[s.np][0] = pt[0];
            s.p[s.np][1] = pt[1];
            s.np += 1;
            if (s.np == 4) panel(&s, s.np);
            stop = cut;
        }
    next:
        rows += 1;
        px = xv;
        py = v;
        if (stop) break;
    }
    if (rows == 0) {
        *value = *rfrom = *rto = NAN;
        return;
    }
    while (s.np > 1) panel(&s, s.np);
    const double f = count > (uint64_t)(cut ? 1 : 0) ? first_x : NAN;
    const double t = cut ? cut_x : px;
    const double total = s.sum[0] + s.sum[1] + s.sum[2];
    *value = rms ? sqrt(total / (t - f)) : total;
    *rfrom = f;
    *rto = t;
}

/* FIND (slope 0) or DERIV (1) of y2 WHEN y crosses as `cl` says: the value
 * at the WHEN's event (measure.zig's Wave.valueAt / slopeAt there). The
 * device's Find (scan.zig:339-380) measures this as it goes, with the same
 * bits, for a WHEN without TD=, FROM= or a LAST count. */
double or_meas_find_when(const double *x, const double *y, const double *y2, uint32_t len, const or_meas_clause *cl,
                         int slope)
{
    const double at = or_meas_cross(x, y, len, cl);
    if (isnan(at)) return NAN;
    return or_meas_at(x, y2, len, at, slope);
}
