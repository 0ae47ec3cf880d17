"""GPU parity of the transient lanes against the CPU oracle, through the C ABI.

Each lane runs the whole transient on the GPU (operating point, steps,
breakpoints, LTE control, .meas passes); the oracle runs the same lane on
the CPU. Every recorded row, every step statistic and every .meas value
must match bit for bit (tolerance: 0 ulp; a NaN matches any NaN). The
nominal lane of docs/examples/measure.sp must also reproduce the
reference's printed results (docs/learn/measure.md:24-36). Each case prints
its name and sizes before touching the GPU.
"""
import ctypes as C
import math
import sys

import numpy as np
import pytest

import egspice_amd as E
from egspice_amd import workloads as W
import oracle_lib as O

pytestmark = pytest.mark.gpu

CAP = 4096
MAX_STEPS = 200_000


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build_both(oracle, builder, **kw):
    t = E.Template()
    builder(t.add, **kw)
    t.freeze()
    c = O.Circuit(oracle)
    builder(c.add, **kw)
    c.freeze()
    assert (t.col_ptr == c.col_ptr).all() and (t.row_idx == c.row_idx).all()
    return t, c


def bits_equal(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    same = a.view(np.uint64) == b.view(np.uint64)
    return bool(np.all(same | (np.isnan(a) & np.isnan(b))))


def P(a):
    return np.ascontiguousarray(a).ctypes.data_as(C.POINTER(C.c_double))


def printed(value, golden, digits):
    assert value == pytest.approx(golden, rel=0, abs=0.5 * 10 ** (math.floor(math.log10(abs(golden))) - digits) * 1.0001)


def measure_passes(t):
    """The ten .meas lines of docs/examples/measure.sp as lane passes."""
    vin, vout, iv1 = t.node("in"), t.node("out"), 2
    return [
        E.meas_cross(vout, 0.1, rise=1),            # trise TRIG
        E.meas_cross(vout, 0.9, rise=1),            # trise TARG
        E.meas_cross(vin, 0.5, rise=1),             # tpd TRIG
        E.meas_cross(vout, 0.5, rise=1),            # tpd TARG
        E.meas_at(vout, 3e-6),                      # vout_3u FIND AT
        E.meas_cross(vout, 0.5, fall=1),            # t_half WHEN FALL
        E.meas_extremum(vout, 1),                   # vmax MAX
        E.meas_extremum(vout, 2, frm=0.0, to=10e-6),  # vavg AVG
        E.meas_integ(vout, rms=True, frm=0.0, to=10e-6),  # vrms RMS
        E.meas_at(vout, 2e-6, deriv=True),          # slope DERIV AT
        E.meas_integ(iv1, frm=1e-6, to=6e-6),       # q_src INTEG I(V1)
    ]


def oracle_meas(oracle, passes, tt, xx):
    """The oracle's scan passes over one recorded waveform, shaped like
    the engine's (value, aux0, aux1) rows."""
    N = len(tt)
    out = np.full((len(passes), 3), np.nan)
    for k, p in enumerate(passes):
        y = np.ascontiguousarray(xx[:, p.row])
        if p.kind == E.MEAS_CROSS:
            cl = O.clause(val=p.val, rise=p.rise, fall=p.fall, cross=p.cross, td=p.td, frm=p.frm, to=p.to)
            out[k, 0] = oracle.or_meas_cross(P(tt), P(y), N, C.byref(cl))
        elif p.kind == E.MEAS_AT:
            out[k, 0] = oracle.or_meas_at(P(tt), P(y), N, p.at, p.op)
        elif p.kind == E.MEAS_FIND:
            w = passes[k - 1]
            yw = np.ascontiguousarray(xx[:, w.row])
            cl = O.clause(val=w.val, rise=w.rise, fall=w.fall, cross=w.cross, td=w.td, frm=w.frm, to=w.to)
            out[k, 0] = oracle.or_meas_find_when(P(tt), P(yw), P(y), N, C.byref(cl), p.op)
        elif p.kind == E.MEAS_EXTREMUM:
            v, a = C.c_double(), C.c_double()
            oracle.or_meas_extremum(P(tt), P(y), N, p.frm, p.to, p.op, C.byref(v), C.byref(a))
            out[k, :2] = v.value, a.value
        else:
            v, a, b = C.c_double(), C.c_double(), C.c_double()
            oracle.or_meas_integ(P(tt), P(y), N, p.frm, p.to, p.op, C.byref(v), C.byref(a), C.byref(b))
            out[k] = v.value, a.value, b.value
    return out


def run_pair(oracle, t, c, sets, opts_kw, passes=(), check_meas=True):
    """Engine lanes vs oracle lanes; returns the engine's recording."""
    N = len(next(iter(sets.values()))) if sets else 1
    gl = E.Lanes(t, N)
    ol = O.Lanes(c, N)
    for (kind, inst, k), vals in sets.items():
        gl.set_param(kind, inst, k, vals)
        ol.set_param(kind, inst, k, vals)
    gl.prepare()
    ol.prepare()
    go = E.tran_opts(**opts_kw, max_steps=MAX_STEPS, record_capacity=CAP)
    res, meas, times, rows, ms = gl.tran(go, passes)
    oo = O.tran_opts(**opts_kw, max_steps=MAX_STEPS)
    say(f"  kernel {ms:.3f} ms, points {res['points'].min()}..{res['points'].max()}, "
        f"attempts {res['attempts'].min()}..{res['attempts'].max()}")
    for l in range(N):
        tt, xx, st = ol.tran(l, oo)
        k = int(res["points"][l])
        assert st.status == res["status"][l], l
        assert st.points == k, (l, st.points, k)
        assert k <= CAP
        assert (st.steps, st.attempts, st.rej_newton, st.rej_lte, st.bp_landings, st.nr_iters) == \
            tuple(int(res[f][l]) for f in ("steps", "attempts", "rej_newton", "rej_lte", "bp_landings", "nr_iters")), l
        assert bits_equal(times[l, :k], tt), l
        assert bits_equal(rows[l, :k], xx), l
        if passes and check_meas:
            assert bits_equal(meas[l], oracle_meas(oracle, passes, tt, xx)), (l, meas[l])
    return res, meas, times, rows


def test_measure_sp_lanes_bitwise(oracle, engine):
    N = 64
    say(f"case: measure.sp RC, {N} lanes (lane 0 nominal, R and C varied), tran 10n 10u, 11 .meas passes")
    t, c = build_both(oracle, W.rc_pulse)
    rng = np.random.default_rng(11)
    r = rng.uniform(500.0, 2000.0, N)
    cap = rng.uniform(0.5e-9, 2e-9, N)
    r[0], cap[0] = 1e3, 1e-9
    passes = measure_passes(t)
    res, meas, times, _ = run_pair(oracle, t, c, {(E.RESISTOR, 0, 0): r, (E.CAPACITOR, 0, 0): cap},
                                   dict(t_stop=10e-6, dt_init=10e-9), passes)
    assert (res["status"] == 0).all()
    # the reference's printed results on the nominal lane
    m = meas[0, :, 0]
    assert res["points"][0] == 1029
    printed(m[1] - m[0], 2.197201e-06, 6)
    printed(m[1], 3.303064e-06, 6)
    printed(m[0], 1.105863e-06, 6)
    printed(m[3] - m[2], 6.931367e-07, 6)
    printed(m[3], 1.693637e-06, 6)
    printed(m[2], 1.000500e-06, 6)
    printed((m[1] - m[0]) / (m[3] - m[2]), 3.169938e+00, 6)
    printed(m[4], 8.645984e-01, 6)
    printed(m[5], 6.68790e-06, 5)
    printed(m[6], 9.932657e-01, 6)
    printed(meas[0, 6, 1], 6.001000e-06, 6)
    printed(m[7], 4.982781e-01, 6)
    printed(m[8], 6.33054e-01, 5)
    printed(m[9], 3.685512e+05, 6)
    printed(m[10], -9.93246e-10, 5)
    assert tuple(meas[0, 10, 1:]) == (1e-6, 6e-6)


def test_find_when_lanes_bitwise(oracle, engine):
    """FIND / DERIV ... WHEN as device passes (the WHEN's Cross, then the
    Find reading it) on measure.sp's lanes, bit for bit against the oracle's
    value at the WHEN's event; an event that never comes gives NaN on both
    sides, and a Find whose WHEN has TD= is refused."""
    N = 32
    say(f"case: measure.sp RC, {N} lanes, FIND/DERIV ... WHEN passes")
    t, c = build_both(oracle, W.rc_pulse)
    rng = np.random.default_rng(12)
    r = rng.uniform(500.0, 2000.0, N)
    cap = rng.uniform(0.5e-9, 2e-9, N)
    out, vin = t.node("out"), t.node("in")
    passes = [E.meas_cross(out, 0.5, fall=1), E.meas_find(vin),
              E.meas_cross(vin, 0.5, rise=1), E.meas_find(out, deriv=True),
              E.meas_cross(out, 0.3, rise=2), E.meas_find(out),
              E.meas_cross(out, 0.7, cross=1, to=8e-6), E.meas_find(out, deriv=True)]
    res, meas, _, _ = run_pair(oracle, t, c, {(E.RESISTOR, 0, 0): r, (E.CAPACITOR, 0, 0): cap},
                               dict(t_stop=10e-6, dt_init=10e-9), passes)
    assert (res["status"] == 0).all()
    assert np.all(meas[:, 1, 0] == 0.0)                     # v(in) is back at 0 when out falls
    assert np.all(np.isfinite(meas[:, 3, 0]))
    assert np.all(np.isnan(meas[:, 5, 0]))                  # a second rise never comes
    lanes = E.Lanes(t, 1)
    lanes.prepare()
    with pytest.raises(E.EngineError):
        lanes.tran(E.tran_opts(10e-6, 10e-9), [E.meas_cross(out, 0.5, rise=1, td=1e-6), E.meas_find(vin)])


def test_rc_lowpass_lanes_bitwise(oracle, engine):
    N = 8
    say(f"case: rc_lowpass.sp, {N} lanes, tran 10u 5m")
    t, c = build_both(oracle, W.rc_pulse, c=159e-9, pulse=W.RC_LOWPASS_PULSE, dc=0.0)
    r = 1e3 * (1 + 0.05 * np.arange(N))
    res, _, _, _ = run_pair(oracle, t, c, {(E.RESISTOR, 0, 0): r}, dict(t_stop=5e-3, dt_init=10e-6))
    assert res["points"][0] == 519
    assert (res["status"] == 0).all()


@pytest.mark.parametrize("method", [0, 1, 2])
def test_rc_ladder_methods_bitwise(oracle, engine, method):
    N = 32
    say(f"case: 4-stage RC ladder, {N} lanes, method {method}, tran 10n 10u")
    t, c = build_both(oracle, W.rc_ladder, stages=4)
    rng = np.random.default_rng(100 + method)
    sets = {(E.RESISTOR, k, 0): rng.uniform(300.0, 3000.0, N) for k in range(4)}
    sets[(E.CAPACITOR, 3, 0)] = rng.uniform(0.2e-9, 3e-9, N)
    passes = [E.meas_cross(t.node("n4"), 0.5, rise=1), E.meas_cross(t.node("n4"), 0.5, fall=E.LAST),
              E.meas_extremum(t.node("n2"), 0), E.meas_integ(t.node("n3"), rms=True)]
    res, _, _, _ = run_pair(oracle, t, c, sets, dict(t_stop=10e-6, dt_init=10e-9, method=method), passes)
    assert (res["status"] == 0).all()


def test_tran_many_lanes_bitwise_sample(oracle, engine):
    """4096 lanes in one launch; a sample of them checked against the
    oracle (the rest only for completion)."""
    N = 4096
    say(f"case: measure.sp RC, {N} lanes, sampled parity")
    t, c = build_both(oracle, W.rc_pulse)
    rng = np.random.default_rng(7)
    r = rng.uniform(500.0, 2000.0, N)
    cap = rng.uniform(0.5e-9, 2e-9, N)
    gl = E.Lanes(t, N)
    gl.set_param(E.RESISTOR, 0, 0, r)
    gl.set_param(E.CAPACITOR, 0, 0, cap)
    gl.prepare()
    passes = measure_passes(t)
    res, meas, times, rows, ms = gl.tran(E.tran_opts(10e-6, 10e-9, max_steps=MAX_STEPS, record_capacity=0), passes)
    say(f"  kernel {ms:.3f} ms for {N} lanes")
    assert times is None and (res["status"] == 0).all()
    sample = [0, 1, 2047, 4095]
    ol = O.Lanes(c, len(sample))
    ol.set_param(O.RES, 0, 0, r[sample])
    ol.set_param(O.CAP, 0, 0, cap[sample])
    ol.prepare()
    for j, l in enumerate(sample):
        tt, xx, st = ol.tran(j, O.tran_opts(10e-6, 10e-9, max_steps=MAX_STEPS))
        assert st.points == res["points"][l]
        assert bits_equal(meas[l], oracle_meas(oracle, passes, tt, xx)), l


def test_cmos_inverter_lanes_bitwise(oracle, engine):
    """docs/examples/cmos_inverter.sp (mos1, no oxide or junction charge):
    lane 0 nominal must give the reference's 2032 points and tphl; the other
    lanes vary VDD, the NMOS threshold and CL."""
    N = 48
    say(f"case: cmos_inverter.sp mos1, {N} lanes, tran 10p 20n")
    t, c = build_both(oracle, W.cmos_inverter)
    rng = np.random.default_rng(5)
    vdd = rng.uniform(1.6, 2.0, N)
    vto = 0.45 + 0.015 * rng.standard_normal(N)
    cl = rng.uniform(5e-15, 20e-15, N)
    vdd[0], vto[0], cl[0] = 1.8, 0.45, 10e-15
    M = E.MOS1_PARAMS
    passes = [E.meas_cross(t.node("in"), 0.9, rise=1), E.meas_cross(t.node("out"), 0.9, fall=1)]
    res, meas, _, _ = run_pair(oracle, t, c, {(E.VSOURCE, 0, 0): vdd, (E.MOS1, 0, M["VTO"]): vto,
                                              (E.CAPACITOR, 0, 0): cl}, dict(t_stop=20e-9, dt_init=10e-12), passes)
    assert (res["status"] == 0).all()
    assert res["points"][0] == 2032
    trig, targ = meas[0, 0, 0], meas[0, 1, 0]
    printed(targ - trig, 2.818778e-11, 6)
    printed(targ, 1.078188e-09, 6)
    printed(trig, 1.050000e-09, 6)


def test_diode_rectifier_lanes_bitwise(oracle, engine):
    """docs/examples/diode_rectifier.sp: SIN source, a diode with RS (its
    junction on an internal row) and depletion charge. Lane 0 nominal must
    give the reference's 610 points, vpeak and vmin (docs/learn/models.md:
    38-45); the other lanes vary RS, CJO, TT (diffusion charge), M and the
    load."""
    N = 24
    say(f"case: diode_rectifier.sp, {N} lanes, tran 0.1m 60m, MAX and MIN passes")
    t, c = build_both(oracle, W.diode_rectifier)
    rng = np.random.default_rng(21)
    D = E.DIODE_PARAMS
    rs = rng.uniform(0.2, 2.0, N)
    cjo = 10 ** rng.uniform(-12, -10, N)
    tt = np.where(np.arange(N) % 3 == 0, 0.0, 10 ** rng.uniform(-9, -6, N))
    mj = rng.uniform(0.3, 0.6, N)
    rl = rng.uniform(500.0, 2000.0, N)
    rs[0], cjo[0], tt[0], mj[0], rl[0] = 0.5, 10e-12, 0.0, 0.5, 1e3
    vout = t.node("out")
    passes = [E.meas_extremum(vout, 1), E.meas_extremum(vout, 0, frm=40e-3, to=60e-3)]
    sets = {(E.DIODE, 0, D["RS"]): rs, (E.DIODE, 0, D["CJO"]): cjo, (E.DIODE, 0, D["TT"]): tt,
            (E.DIODE, 0, D["M"]): mj, (E.RESISTOR, 0, 0): rl}
    res, meas, _, _ = run_pair(oracle, t, c, sets, dict(t_stop=60e-3, dt_init=0.1e-3), passes)
    assert (res["status"] == 0).all()
    assert res["points"][0] == 610
    printed(meas[0, 0, 0], 9.228482e+00, 6)
    printed(meas[0, 0, 1], 5.173860e-03, 6)
    printed(meas[0, 1, 0], 7.719858e+00, 6)
    printed(meas[0, 1, 1], 4.317386e-02, 6)


def test_collapsed_diode_charge_lanes_bitwise(oracle, engine):
    """A diode without RS (ai collapsed onto the anode) carrying depletion
    and diffusion charge, driven by PULSE edges through a resistor: reverse
    recovery stresses the charge's limiting correction and the LTE site."""
    N = 16
    say(f"case: collapsed diode with CJO and TT, {N} lanes, tran 1n 200n")

    def deck(add):
        add(E.VSOURCE, ["in", "0"], W.pulse_params(-2.0, 2.0, 10e-9, 1e-9, 1e-9, 40e-9, 100e-9))
        add(E.DIODE, ["a", "0"], {E.DIODE_PARAMS["IS"]: 1e-14, E.DIODE_PARAMS["CJO"]: 2e-12,
                                  E.DIODE_PARAMS["TT"]: 5e-9})
        add(E.RESISTOR, ["in", "a"], {0: 1e3})

    t, c = build_both(oracle, deck)
    assert t.n == c.n == 4  # in, V's branch, a: no internal row
    rng = np.random.default_rng(33)
    D = E.DIODE_PARAMS
    sets = {(E.DIODE, 0, D["CJO"]): 10 ** rng.uniform(-13, -11, N), (E.DIODE, 0, D["TT"]): 10 ** rng.uniform(-10, -8, N),
            (E.DIODE, 0, D["VJ"]): rng.uniform(0.6, 0.9, N), (E.RESISTOR, 0, 0): rng.uniform(200.0, 2000.0, N)}
    passes = [E.meas_cross(t.node("a"), 0.0, fall=1), E.meas_extremum(t.node("a"), 0)]
    for method in (0, 1, 2):
        res, _, _, _ = run_pair(oracle, t, c, sets, dict(t_stop=200e-9, dt_init=1e-9, method=method), passes)
        assert (res["status"] == 0).all()


def test_pwl_lanes_bitwise(oracle, engine):
    """A PWL source (4 points, with r= repeat and td= delay per lane) into
    the RC of measure.sp: its corners land as ngspice rounds them, its
    repeats fold; every row bitwise the oracle's."""
    N = 16
    say(f"case: PWL RC, {N} lanes (r= and td= varied), tran 10n 4u")

    def deck(add):
        p = {0: np.nan, 1: 4.0, 15: 4.0}
        for k, (tk, vk) in enumerate(zip([0.0, 1e-7, 3e-7, 6e-7], [0.0, 1.0, 1.0, -0.5])):
            p[18 + k], p[34 + k] = tk, vk
        add(E.VSOURCE, ["in", "0"], p)
        add(E.CAPACITOR, ["out", "0"], {0: 1e-9})
        add(E.RESISTOR, ["in", "out"], {0: 1e3})

    t, c = build_both(oracle, deck)
    rng = np.random.default_rng(17)
    rep = np.where(np.arange(N) % 2 == 0, -1.0, rng.uniform(0.0, 5e-7, N))
    td = np.where(np.arange(N) % 3 == 0, 0.0, rng.uniform(0.0, 2e-7, N))
    sets = {(E.VSOURCE, 0, 16): rep, (E.VSOURCE, 0, 17): td, (E.RESISTOR, 0, 0): rng.uniform(200.0, 2000.0, N)}
    passes = [E.meas_cross(t.node("out"), 0.25, rise=1), E.meas_extremum(t.node("out"), 0)]
    res, _, _, _ = run_pair(oracle, t, c, sets, dict(t_stop=4e-6, dt_init=1e-8), passes)
    assert (res["status"] == 0).all()


def test_exp_lanes_bitwise(oracle, engine):
    """An EXP source (models/vsource.va:252-258) into measure.sp's RC, its
    delays and time constants varied per lane, some left at 0 so the
    .tran step fills them (frontend/builder.zig:862-867); every row bitwise
    the oracle's."""
    N = 16
    say(f"case: EXP RC, {N} lanes (td1 tau1 td2 tau2 varied, zeros filled from the step), tran 10n 4u")

    def deck(add):
        add(E.VSOURCE, ["in", "0"], {0: np.nan, 1: 3.0, 50: 0.0, 51: 1.0, 52: 2e-7, 53: 1e-7, 54: 2e-6, 55: 3e-7})
        add(E.CAPACITOR, ["out", "0"], {0: 1e-9})
        add(E.RESISTOR, ["in", "out"], {0: 1e3})

    t, c = build_both(oracle, deck)
    rng = np.random.default_rng(23)
    zero = lambda k: np.where(np.arange(N) % k == 0, 0.0, 1.0)
    sets = {(E.VSOURCE, 0, 52): zero(4) * rng.uniform(0.0, 1e-6, N),
            (E.VSOURCE, 0, 53): zero(5) * rng.uniform(2e-8, 5e-7, N),
            (E.VSOURCE, 0, 54): zero(3) * rng.uniform(1e-6, 3e-6, N),
            (E.VSOURCE, 0, 55): zero(7) * rng.uniform(2e-8, 5e-7, N),
            (E.VSOURCE, 0, 51): rng.uniform(-1.0, 2.0, N)}
    passes = [E.meas_cross(t.node("out"), 0.25, rise=1), E.meas_extremum(t.node("out"), 1)]
    res, _, _, _ = run_pair(oracle, t, c, sets, dict(t_stop=4e-6, dt_init=1e-8), passes)
    assert (res["status"] == 0).all()


def test_isource_waveform_lanes_bitwise(oracle, engine):
    """Current sources with PULSE and EXP waveforms (models/isource.va: the
    voltage source's slots, expressions and timers) into an RC, their
    levels and times varied per lane; every row bitwise the oracle's."""
    N = 16
    say(f"case: I-source PULSE + EXP into RC, {N} lanes, tran 10n 4u")

    def deck(add):
        add(E.ISOURCE, ["0", "in"], {0: np.nan, 1: 1.0, 2: 0.0, 3: 1e-3, 4: 3e-7, 5: 1e-7, 6: 2e-7, 7: 1e-6,
                                     8: 2.5e-6})
        add(E.ISOURCE, ["0", "in"], {0: np.nan, 1: 3.0, 50: 0.0, 51: 5e-4, 52: 2e-7, 53: 1e-7, 54: 2e-6})
        add(E.CAPACITOR, ["out", "0"], {0: 1e-9})
        add(E.RESISTOR, ["in", "out"], {0: 1e3})
        add(E.RESISTOR, ["in", "0"], {0: 1e3})

    t, c = build_both(oracle, deck)
    rng = np.random.default_rng(29)
    sets = {(E.ISOURCE, 0, 3): rng.uniform(2e-4, 2e-3, N), (E.ISOURCE, 0, 4): rng.uniform(0.0, 1e-6, N),
            (E.ISOURCE, 1, 53): np.where(np.arange(N) % 3 == 0, 0.0, rng.uniform(5e-8, 5e-7, N)),
            (E.RESISTOR, 0, 0): rng.uniform(200.0, 2000.0, N)}
    passes = [E.meas_cross(t.node("out"), 0.25, rise=1), E.meas_extremum(t.node("out"), 1)]
    res, _, _, _ = run_pair(oracle, t, c, sets, dict(t_stop=4e-6, dt_init=1e-8), passes)
    assert (res["status"] == 0).all()


def test_sffm_am_lanes_bitwise(oracle, engine):
    """SFFM (FM and FC from the stop time when omitted) and AM voltage
    sources into an RC (models/vsource.va:304-330), their frequencies,
    indices and delays varied per lane; every row bitwise the oracle's."""
    N = 16
    say(f"case: SFFM + AM sources into RC, {N} lanes, tran 10n 4u")

    def deck(add):
        add(E.VSOURCE, ["f", "0"], {0: np.nan, 1: 5.0, 56: 0.1, 57: 1.0, 59: 2.0, 61: 2e-7})
        add(E.VSOURCE, ["a", "0"], {0: np.nan, 1: 6.0, 64: 1.5, 65: 0.5, 66: 1e6, 67: 5e6, 68: 5e-7})
        add(E.CAPACITOR, ["out", "0"], {0: 1e-10})
        add(E.RESISTOR, ["f", "out"], {0: 1e3})
        add(E.RESISTOR, ["a", "out"], {0: 2e3})

    t, c = build_both(oracle, deck)
    rng = np.random.default_rng(37)
    sets = {(E.VSOURCE, 0, 58): np.where(np.arange(N) % 2 == 0, -1.0, rng.uniform(2e5, 2e6, N)),
            (E.VSOURCE, 0, 59): rng.uniform(0.0, 8.0, N),
            (E.VSOURCE, 1, 66): rng.uniform(5e5, 3e6, N), (E.VSOURCE, 1, 68): rng.uniform(0.0, 1e-6, N)}
    passes = [E.meas_extremum(t.node("out"), 1), E.meas_integ(t.node("out"), rms=True)]
    res, _, _, _ = run_pair(oracle, t, c, sets, dict(t_stop=4e-6, dt_init=1e-8), passes)
    assert (res["status"] == 0).all()


def test_isource_sffm_is_refused(engine):
    say("case: SFFM on a current source is refused at prepare")
    t = E.Template()
    t.add(E.ISOURCE, ["0", "a"], {0: np.nan, 1: 5.0, 56: 0.0, 57: 1e-3})
    t.add(E.RESISTOR, ["a", "0"], {0: 1e3})
    t.freeze()
    gl = E.Lanes(t, 2)
    with pytest.raises(E.EngineError, match="IsourceSffmUnsupported"):
        gl.prepare()


def test_diode_rs_topology_is_per_template(engine):
    say("case: a lane may not move a diode's RS across zero")
    t = E.Template()
    W.diode_rectifier(t.add)
    t.freeze()
    gl = E.Lanes(t, 2)
    gl.set_param(E.DIODE, 0, E.DIODE_PARAMS["RS"], [0.5, 0.0])
    with pytest.raises(E.EngineError, match="DiodeRsTopology"):
        gl.prepare()


def test_tran_refuses_charged_models(engine):
    say("case: a MOSFET with oxide charge is refused by the transient lanes")
    M = E.MOS1_PARAMS
    t = E.Template()
    t.add(E.VSOURCE, ["d", "0"], {0: 1.8})
    t.add(E.VSOURCE, ["g", "0"], {0: 1.0})
    t.add(E.RESISTOR, ["d", "o"], {0: 1e4})
    t.add(E.MOS1, ["o", "g", "0", "0"], {**W.INV_NCH, M["W"]: 1e-6, M["L"]: 180e-9})
    t.freeze()
    gl = E.Lanes(t, 2)
    gl.set_param(E.MOS1, 0, M["TOX"], [float("nan"), 10e-9])
    gl.prepare()
    with pytest.raises(E.EngineError, match="TranChargesUnsupported"):
        gl.tran(E.tran_opts(1e-6, 1e-8))
    gl.set_param(E.MOS1, 0, M["TOX"], [float("nan"), float("nan")])
    gl.prepare()
    res, _, _, _, _ = gl.tran(E.tran_opts(1e-6, 1e-8))
    assert (res["status"] == 0).all()
