"""The points API (include/espice.h) and the espice CLI on the GPU.

The documented decks run through espice_create / espice_append_points /
espice_run_all / espice_copy_point_measures, and every measure must equal,
bit for bit, the oracle's transient and .meas passes on the same circuit with
the deck's numbers read the same way (ngspice's INPevaluate, so `10u` is
10 * 1e-6). The nominal points must also reproduce the reference's printed
results (docs/learn/measure.md:24-36, models.md:95-102), and the CLI must
print them as the reference does. Each case prints its name before touching
the GPU.
"""
import ctypes as C
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import egspice_amd as E
from egspice_amd import workloads as W
import oracle_lib as O
import decks

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sn = E.spice_number


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def P(a):
    return np.ascontiguousarray(a).ctypes.data_as(C.POINTER(C.c_double))


def printed(value, golden, digits):
    assert value == pytest.approx(golden, rel=0, abs=0.5 * 10 ** (math.floor(math.log10(abs(golden))) - digits) * 1.0001)


def bits_equal(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    same = a.view(np.uint64) == b.view(np.uint64)
    return bool(np.all(same | (np.isnan(a) & np.isnan(b))))


MEAS_NAMES = ["trise", "tdelay", "vat3u", "tcross", "vmax", "vavg", "vrms", "slope", "charge", "ratio"]


def oracle_measure_sp(oracle, r, c):
    """measure.sp on the oracle with the deck's numbers as INPevaluate reads
    them; the ten .meas values."""
    pulse = tuple(sn(x) for x in ("0", "1", "1u", "1n", "1n", "5u", "20u"))
    ck = O.Circuit(oracle)
    W.rc_pulse(ck.add, r=r, c=c, pulse=pulse)
    ck.freeze()
    L = O.Lanes(ck, 1)
    L.prepare()
    t, x, st = L.tran(0, O.tran_opts(sn("10u"), sn("10n"), max_steps=200_000))
    assert st.status == 0
    n = len(t)
    vin, vout, iv = (np.ascontiguousarray(x[:, k]) for k in (ck.node("in"), ck.node("out"), 2))
    cross = lambda y, **k: oracle.or_meas_cross(P(t), P(y), n, C.byref(O.clause(**k)))
    trise = cross(vout, val=sn("0.9"), rise=1) - cross(vout, val=sn("0.1"), rise=1)
    tdelay = cross(vout, val=sn("0.5"), rise=1) - cross(vin, val=sn("0.5"), rise=1)
    v, a, b = C.c_double(), C.c_double(), C.c_double()
    oracle.or_meas_extremum(P(t), P(vout), n, 0.0, 0.0, 1, C.byref(v), C.byref(a))
    vmax = v.value
    oracle.or_meas_extremum(P(t), P(vout), n, sn("0"), sn("10u"), 2, C.byref(v), C.byref(a))
    vavg = v.value
    oracle.or_meas_integ(P(t), P(vout), n, sn("0"), sn("10u"), 1, C.byref(v), C.byref(a), C.byref(b))
    vrms = v.value
    oracle.or_meas_integ(P(t), P(iv), n, sn("1u"), sn("6u"), 0, C.byref(v), C.byref(a), C.byref(b))
    charge = v.value
    return [trise, tdelay, oracle.or_meas_at(P(t), P(vout), n, sn("3u"), 0), cross(vout, val=sn("0.5"), fall=1),
            vmax, vavg, vrms, oracle.or_meas_at(P(t), P(vout), n, sn("2u"), 1), charge, trise / tdelay], len(t)


def test_measure_sp_nominal_point(oracle, engine):
    say("case: measure.sp through espice_create/run_all/copy_point_measures, point 0")
    pb = E.Problem(decks.MEASURE_SP)
    assert pb.query_count() == 1 and pb.lane_width() == 16384
    pb.run_all()
    m, reasons = pb.measures(0, 1, MEAS_NAMES)
    assert reasons.tolist() == [0]
    for k, name in enumerate(MEAS_NAMES):
        printed(m[0, k], *decks.MEASURE_GOLDEN[name])
    want, points = oracle_measure_sp(oracle, sn("1k"), sn("1n"))
    assert points == decks.MEASURE_POINTS
    assert bits_equal(m[0], want)
    assert "Transient Analysis: 1029 points, 4 variables" in pb.report()


def test_find_when_cards(oracle, engine):
    """`.meas tran ... FIND/DERIV ... WHEN` cards through the deck reader and
    the points API: the WHEN's Cross and the Find as device passes, bitwise
    the oracle's value at the WHEN's event on the same waveform."""
    say("case: measure.sp with FIND ... WHEN and DERIV ... WHEN cards, point 0")
    text = decks.MEASURE_SP.replace(".end", ".meas tran vin_at FIND v(in) WHEN v(out)=0.5 FALL=1\n"
                                           ".meas tran sl_at DERIV v(out) WHEN v(out)=0.3 RISE=1\n"
                                           ".meas tran never FIND v(out) WHEN v(out)=0.3 RISE=2\n.end")
    pb = E.Problem(text)
    pb.run_all()
    m, reasons = pb.measures(0, 1, ["vin_at", "sl_at", "never"])
    pulse = tuple(sn(x) for x in ("0", "1", "1u", "1n", "1n", "5u", "20u"))
    ck = O.Circuit(oracle)
    W.rc_pulse(ck.add, r=sn("1k"), c=sn("1n"), pulse=pulse)
    ck.freeze()
    L = O.Lanes(ck, 1)
    L.prepare()
    t, x, st = L.tran(0, O.tran_opts(sn("10u"), sn("10n"), max_steps=200_000))
    n = len(t)
    vin, vout = (np.ascontiguousarray(x[:, k]) for k in (ck.node("in"), ck.node("out")))
    fw = lambda y, y2, slope, **k: oracle.or_meas_find_when(P(t), P(y), P(y2), n, C.byref(O.clause(**k)), slope)
    want = [fw(vout, vin, 0, val=sn("0.5"), fall=1), fw(vout, vout, 1, val=sn("0.3"), rise=1),
            fw(vout, vout, 0, val=sn("0.3"), rise=2)]
    assert bits_equal(m[0], want)
    assert m[0, 0] == 0.0 and np.isfinite(m[0, 1]) and np.isnan(m[0, 2])


def test_measure_sp_live_points(oracle, engine):
    N = 48
    say(f"case: measure.sp with live r, c: {N} points in one lane batch, bitwise vs the oracle")
    pb = E.Problem(decks.MEASURE_PARAM_SP, live=("r", "c"))
    rng = np.random.default_rng(3)
    # tau = RC up to 1.44 us, so v(out) passes 0.9 V within the 5 us pulse
    # and every measure resolves
    rows = np.stack([rng.uniform(500, 1200, N), rng.uniform(0.5e-9, 1.2e-9, N)], axis=1)
    first = pb.append_points(rows)
    assert first == 1 and pb.query_count() == 1 + N
    pb.run_all()
    m, reasons = pb.measures(first, N, MEAS_NAMES)
    assert (reasons == 0).all()
    for l in range(0, N, 6):
        want, _ = oracle_measure_sp(oracle, rows[l, 0], rows[l, 1])
        assert bits_equal(m[l], want), l
    # the nominal point is still point 0, unchanged by the appended ones
    m0, _ = pb.measures(0, 1, ["trise"])
    printed(m0[0, 0], 2.197201e-06, 6)
    # released points give RESULT_UNAVAILABLE, their ids stay known
    pb.release(first, N)
    with pytest.raises(E.PointsError) as e:
        pb.measures(first, N, ["trise"])
    assert e.value.status == 5


def test_point_reasons(engine):
    say("case: measure.sp points with a non-finite value and an unresolved measure")
    pb = E.Problem(decks.MEASURE_PARAM_SP.replace(".end", ".meas tran never WHEN v(out)=5 RISE=1\n.end"),
                   live=("r", "c"))
    first = pb.append_points([[1e3, 1e-9], [float("nan"), 1e-9]])
    pb.run_all()
    m, reasons = pb.measures(first, 2, ["trise", "never"])
    assert reasons.tolist() == [5, 9]  # MEASURE_UNRESOLVED, INVALID_PARAMETER
    assert np.isfinite(m[0, 0]) and np.isnan(m[0, 1]) and np.isnan(m[1]).all()
    with pytest.raises(E.PointsError, match="UnknownMeasure"):
        pb.measures(first, 2, ["nope"])
    with pytest.raises(E.PointsError) as e:
        pb.measures(first + 1, 2, ["trise"])  # not a point's first query / past the end
    assert e.value.status == 4


def test_cmos_inverter_points(oracle, engine):
    N = 24
    say(f"case: cmos_inverter.sp, live vdd and wn, {N} points + nominal, bitwise vs the oracle")
    pb = E.Problem(decks.CMOS_INVERTER_SP, live=("vdd", "wn"))
    rng = np.random.default_rng(9)
    rows = np.stack([rng.uniform(1.6, 2.0, N), rng.uniform(0.8e-6, 1.5e-6, N)], axis=1)
    first = pb.append_points(rows)
    pb.run_all()
    m0, r0 = pb.measures(0, 1, ["tphl"])
    assert r0.tolist() == [0]
    printed(m0[0, 0], *decks.CMOS_INVERTER_GOLDEN["tphl"])
    assert f"Transient Analysis: {decks.CMOS_INVERTER_POINTS} points" in pb.report()
    m, reasons = pb.measures(first, N, ["tphl"])
    assert (reasons == 0).all()
    M = E.MOS1_PARAMS
    model = lambda typ, vto, kp: {M["TYPE"]: typ, M["VTO"]: sn(vto), M["KP"]: sn(kp), M["LAMBDA"]: sn("0.05"),
                                  M["GAMMA"]: sn("0.4"), M["PHI"]: sn("0.8"), M["L"]: sn("180n")}
    for l in range(0, N, 4):
        vdd, wn = rows[l]
        ck = O.Circuit(oracle)
        ck.add(O.VSRC, ["vdd", "0"], {0: vdd})
        ck.add(O.VSRC, ["in", "0"], W.pulse_params(sn("0"), vdd, sn("1n"), sn("100p"), sn("100p"), sn("5n"),
                                                   sn("10n"), dc=sn("0")))
        ck.add(O.CAP, ["out", "0"], {0: sn("10f")})
        ck.add(O.MOS1, ["out", "in", "0", "0"], {**model(1.0, "0.45", "200u"), M["W"]: wn})
        ck.add(O.MOS1, ["out", "in", "vdd", "vdd"], {**model(-1.0, "-0.45", "80u"), M["W"]: 2 * wn})
        ck.freeze()
        L = O.Lanes(ck, 1)
        L.prepare()
        t, x, st = L.tran(0, O.tran_opts(sn("20n"), sn("10p"), max_steps=200_000))
        assert st.status == 0
        n = len(t)
        vin, vout = (np.ascontiguousarray(x[:, ck.node(k)]) for k in ("in", "out"))
        trig = oracle.or_meas_cross(P(t), P(vin), n, C.byref(O.clause(val=sn("0.9"), rise=1)))
        targ = oracle.or_meas_cross(P(t), P(vout), n, C.byref(O.clause(val=sn("0.9"), fall=1)))
        assert bits_equal(m[l, 0], targ - trig), l


# src/tests/espice.zig:1789-1798, the reference's measure-bound deck
ABORT_SP = """abort bound
.param rv=1k
v1 in 0 pulse(0 1 0.1n 0.1n 0.1n 20n 40n)
r1 in out {rv}
c1 out 0 0.1p
.tran 1p 10n
.meas tran tpd trig v(in) val=0.5 rise=1 targ v(out) val=0.5 rise=1
.end
"""


def test_points_abort_after_and_early_stop(engine):
    """The reference's own points test (src/tests/espice.zig:1820-1861):
    rv = 1k, 100k, 2k appended twice, plain and with measures_only,
    early_stop and abort_after = 1 ns: the bounded run aborts only the
    100k point (its delay is past the bound), and the other two keep the
    plain run's values bit for bit."""
    say("case: abort.sp points, plain vs measures_only + early_stop + abort_after=1n")
    pb = E.Problem(ABORT_SP, live=("rv",))
    rows = [[1e3], [100e3], [2e3]]
    plain = pb.append_points(rows)
    bounded = pb.append_points(rows, E.points_options(measures_only=True, early_stop=True, abort_after=1e-9))
    pb.run_all()
    want, want_why = pb.measures(plain, 3, ["tpd"])
    got, why = pb.measures(bounded, 3, ["tpd"])
    assert want_why.tolist() == [0, 0, 0]
    assert why.tolist() == [0, 6, 0]  # ok, ESPICE_REASON_ABORTED, ok
    assert np.isnan(got[1, 0])
    assert bits_equal(got[[0, 2], 0], want[[0, 2], 0])
    # a bound on another card than the unresolved one aborts nothing; an
    # unknown card name is refused
    with pytest.raises(E.PointsError, match="UnknownMeasure"):
        pb.append_points(rows, E.points_options(abort_measure="nope", abort_after=1e-9))


def test_points_warm_start(oracle, engine):
    """warm_start: every point's operating point from the template's nominal
    one (op.zig:588, 671): bitwise the oracle's transient started from its
    own nominal operating point, and inside the Newton tolerance of the
    cold start's measures."""
    N = 8
    say(f"case: cmos_inverter.sp, live vdd and wn, {N} points warm from the nominal operating point")
    pb = E.Problem(decks.CMOS_INVERTER_SP, live=("vdd", "wn"))
    rng = np.random.default_rng(19)
    rows = np.stack([rng.uniform(1.6, 2.0, N), rng.uniform(0.8e-6, 1.5e-6, N)], axis=1)
    cold = pb.append_points(rows)
    warm = pb.append_points(rows, E.points_options(warm_start=True))
    pb.run_all()
    mc, rc = pb.measures(cold, N, ["tphl"])
    mw, rw = pb.measures(warm, N, ["tphl"])
    assert (rc == 0).all() and (rw == 0).all()
    assert np.allclose(mw, mc, rtol=1e-3, atol=1e-13)
    M = E.MOS1_PARAMS
    model = lambda typ, vto, kp: {M["TYPE"]: typ, M["VTO"]: sn(vto), M["KP"]: sn(kp), M["LAMBDA"]: sn("0.05"),
                                  M["GAMMA"]: sn("0.4"), M["PHI"]: sn("0.8"), M["L"]: sn("180n")}

    def inverter(vdd, wn):
        ck = O.Circuit(oracle)
        ck.add(O.VSRC, ["vdd", "0"], {0: vdd})
        ck.add(O.VSRC, ["in", "0"], W.pulse_params(sn("0"), vdd, sn("1n"), sn("100p"), sn("100p"), sn("5n"),
                                                   sn("10n"), dc=sn("0")))
        ck.add(O.CAP, ["out", "0"], {0: sn("10f")})
        ck.add(O.MOS1, ["out", "in", "0", "0"], {**model(1.0, "0.45", "200u"), M["W"]: wn})
        ck.add(O.MOS1, ["out", "in", "vdd", "vdd"], {**model(-1.0, "-0.45", "80u"), M["W"]: 2 * wn})
        ck.freeze()
        L = O.Lanes(ck, 1)
        L.prepare()
        return ck, L

    # the template's nominal operating point: the deck's own values, cold
    ck0, L0 = inverter(sn("1.8"), sn("1u"))
    x0, s0, _ = L0.newton_op(mode=1, x0=np.zeros(ck0.n), init_fix=True)
    assert s0[0] & 1
    for l in range(0, N, 3):
        vdd, wn = rows[l]
        ck, L = inverter(vdd, wn)
        t, x, st = L.tran(0, O.tran_opts(sn("20n"), sn("10p"), max_steps=200_000), x0=x0[:, 0])
        assert st.status == 0
        n = len(t)
        vin, vout = (np.ascontiguousarray(x[:, ck.node(k)]) for k in ("in", "out"))
        trig = oracle.or_meas_cross(P(t), P(vin), n, C.byref(O.clause(val=sn("0.9"), rise=1)))
        targ = oracle.or_meas_cross(P(t), P(vout), n, C.byref(O.clause(val=sn("0.9"), fall=1)))
        assert bits_equal(mw[l, 0], targ - trig), l


def test_cli_prints_the_reference_lines(tmp_path, engine):
    say("case: egspice_amd/espice --backend=hip measure.sp")
    deck = tmp_path / "measure.sp"
    deck.write_text(decks.MEASURE_SP)
    exe = os.path.join(ROOT, "egspice_amd", "espice")
    r = subprocess.run([exe, "--backend=hip", "--measures=-", str(deck)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    for want in (
        "trise               =  2.197201e-06 targ=  3.303064e-06 trig=  1.105863e-06",
        "tdelay              =  6.931367e-07 targ=  1.693637e-06 trig=  1.000500e-06",
        "vat3u               =  8.645984e-01",
        "tcross              =   6.68790e-06",
        "vmax                =  9.932657e-01 at=  6.001000e-06",
        "vavg                =  4.982781e-01 from=  0.000000e+00 to=  1.000000e-05",
        "vrms                =   6.33054e-01 from=  0.00000e+00 to=  1.00000e-05",
        "slope               =  3.685512e+05",
        "charge              =   -9.93246e-10 from=  1.00000e-06 to=  6.00000e-06",
        "  Transient Analysis: 1029 points, 4 variables",
    ):
        assert want in lines, (want, r.stdout)
    # PARAM cards print %.5e (src/output/measure.zig:1025-1029)
    assert "ratio               =  3.16994e+00" in lines
    assert lines[-1].startswith("{") and '"trise": 2.19720' in lines[-1]


def test_cli_diode_rectifier(tmp_path, engine):
    say("case: egspice_amd/espice --backend=hip diode_rectifier.sp (SIN, diode RS and CJO)")
    deck = tmp_path / "diode_rectifier.sp"
    deck.write_text(decks.DIODE_RECTIFIER_SP)
    exe = os.path.join(ROOT, "egspice_amd", "espice")
    r = subprocess.run([exe, "--backend=hip", str(deck)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    for want in (
        "vpeak               =  9.228482e+00 at=  5.173860e-03",
        "vmin                =  7.719858e+00 at=  4.317386e-02",
        "Half-wave rectifier: 4 devices",
        "  Transient Analysis: 610 points, 4 variables",
    ):
        assert want in lines, (want, r.stdout)


def test_cli_sources(tmp_path, engine):
    say("case: egspice_amd/espice --backend=hip sources.sp (PULSE, SIN, PWL, EXP, DC current)")
    deck = tmp_path / "sources.sp"
    deck.write_text(decks.SOURCES_SP)
    exe = os.path.join(ROOT, "egspice_amd", "espice")
    r = subprocess.run([exe, "--backend=hip", str(deck)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    for want in ("Independent source waveforms: 10 devices",
                 f"  Transient Analysis: {decks.SOURCES_POINTS} points, 10 variables"):
        assert want in lines, (want, r.stdout)


def test_cli_isources(tmp_path, engine):
    say("case: egspice_amd/espice --backend=hip with I-source PULSE, SIN, PWL, EXP")
    deck = tmp_path / "isources.sp"
    deck.write_text(decks.ISOURCES_SP)
    exe = os.path.join(ROOT, "egspice_amd", "espice")
    r = subprocess.run([exe, "--backend=hip", str(deck)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    for want in ("Independent current source waveforms: 8 devices",
                 f"  Transient Analysis: {decks.SOURCES_POINTS} points, 5 variables"):
        assert want in lines, (want, r.stdout)


def test_diode_rectifier_points(oracle, engine):
    N = 12
    say(f"case: diode_rectifier.sp with live rs, cjo: {N} points, bitwise vs the oracle")
    deck = decks.DIODE_RECTIFIER_SP.replace(".model dmod D(IS=1e-14 N=1.05 RS=0.5 CJO=10p)",
                                            ".param rs=0.5 cj=10p\n.model dmod D(IS=1e-14 N=1.05 RS={rs} CJO={cj})")
    pb = E.Problem(deck, live=("rs", "cj"))
    rng = np.random.default_rng(4)
    rows = np.stack([rng.uniform(0.2, 2.0, N), 10 ** rng.uniform(-12, -10, N)], axis=1)
    first = pb.append_points(rows)
    pb.run_all()
    m0, r0 = pb.measures(0, 1, ["vpeak", "vmin"])
    assert r0.tolist() == [0]
    printed(m0[0, 0], *decks.DIODE_RECTIFIER_GOLDEN["vpeak"])
    printed(m0[0, 1], *decks.DIODE_RECTIFIER_GOLDEN["vmin"])
    m, reasons = pb.measures(first, N, ["vpeak", "vmin"])
    assert (reasons == 0).all()
    D = E.DIODE_PARAMS
    for l in range(0, N, 3):
        ck = O.Circuit(oracle)
        W.diode_rectifier(ck.add, rload=sn("1k"), cres=sn("100u"),
                          dmodel={D["IS"]: sn("1e-14"), D["N"]: sn("1.05"), D["RS"]: rows[l, 0], D["CJO"]: rows[l, 1]})
        ck.freeze()
        L = O.Lanes(ck, 1)
        L.prepare()
        t, x, st = L.tran(0, O.tran_opts(sn("60m"), sn("0.1m"), max_steps=200_000))
        assert st.status == 0
        y = np.ascontiguousarray(x[:, ck.node("out")])
        v, a = C.c_double(), C.c_double()
        oracle.or_meas_extremum(P(t), P(y), len(t), 0.0, 0.0, 1, C.byref(v), C.byref(a))
        vpeak = v.value
        oracle.or_meas_extremum(P(t), P(y), len(t), sn("40m"), sn("60m"), 0, C.byref(v), C.byref(a))
        assert bits_equal(m[l], [vpeak, v.value]), l


def test_cli_refuses_what_the_lanes_cannot_run(tmp_path, engine):
    say("case: espice CLI errors (unsupported card, charged MOSFET)")
    exe = os.path.join(ROOT, "egspice_amd", "espice")
    bad = tmp_path / "bad.sp"
    bad.write_text(decks.MEASURE_SP.replace(".end", "L1 out 0 1u\n.end"))
    r = subprocess.run([exe, "--backend=hip", str(bad)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "UnsupportedElement" in r.stderr
    charged = tmp_path / "charged.sp"
    charged.write_text("mosfet with oxide charge\nV1 d 0 1.8\nV2 g 0 PULSE(0 1.8 1n 1n 1n 5n 10n)\nR1 d o 10k\n"
                       "M1 o g 0 0 nch W=1u L=180n\n.model nch nmos (VTO=0.45 KP=200u TOX=10n)\n.tran 10p 20n\n.end\n")
    r = subprocess.run([exe, "--backend=hip", str(charged)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "TranChargesUnsupported" in r.stderr


def parse_print_table(text):
    """The first plot of a print-format text: its title line, column names
    and rows (up to the next plot's title)."""
    lines = text.splitlines()
    names = lines[1].split("\t")
    rows = []
    for ln in lines[3:]:
        if not ln or not ln[0].isdigit():
            break
        rows.append([float(v) for v in ln.split("\t")])
    return lines[0], names, np.array(rows)


def test_monte_agauss_dc(oracle, engine):
    from test_mc import agauss_trial_op
    say("case: monte_agauss.sp, .dc monte=200 with .param agauss draws, seed 7")
    pb = E.Problem(decks.MONTE_AGAUSS_SP)
    assert pb.query_count() == 201
    pb.run_all()
    head, names, rows = parse_print_table(pb.report(tables=True))
    assert head == "Monte Carlo: Divider spread with agauss and a Monte Carlo DC"
    assert names == ["Index", "run", "i(v1)", "v(in)", "v(out)"]
    assert rows.shape == (200, 5)
    assert (rows[:, 0] == np.arange(200)).all() and (rows[:, 1] == np.arange(1, 201)).all()
    for k, (run, i_v1, v_in, v_out) in enumerate(decks.MONTE_AGAUSS_ROWS):
        assert rows[k, 2] == i_v1 and rows[k, 3] == v_in, k
        assert abs(rows[k, 4] - v_out) <= np.spacing(v_out), k
    for run in (1, 2, 3, 50, 117, 200):
        _, _, x, c = agauss_trial_op(oracle, run)
        assert bits_equal(rows[run - 1, 2:], [x[2], x[c.node("in")], x[c.node("out")]]), run
    # the spread the docs state: sigma(v(out)) about 3.8 mV over 200 trials
    assert 3.0e-3 < rows[:, 4].std() < 4.6e-3


def test_cli_monte_rawfile(tmp_path, engine):
    say("case: espice --rawfile=FILE --format=print monte_agauss.sp")
    deck = tmp_path / "monte_agauss.sp"
    deck.write_text(decks.MONTE_AGAUSS_SP)
    raw = tmp_path / "mc.txt"
    exe = os.path.join(ROOT, "egspice_amd", "espice")
    r = subprocess.run([exe, f"--rawfile={raw}", "--format=print", str(deck)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = raw.read_text().splitlines()
    assert lines[:3] == ["Monte Carlo: Divider spread with agauss and a Monte Carlo DC",
                         "Index\trun\ti(v1)\tv(in)\tv(out)",
                         "-----\t---------------\t---------------\t---------------\t---------------"]
    assert lines[3].startswith("0\t1e0\t-4.98898735103262e-4\t1e0\t")
    assert lines[5] == "2\t3e0\t-5.003152460377415e-4\t1e0\t4.957336298692346e-1"
    assert len(lines) == 203


def test_monte_tran_statistics(engine):
    say("case: monte_tran.sp, PWL source, .tran sweep monte=50 with its WHEN statistics")
    pb = E.Problem(decks.MONTE_TRAN_SP)
    assert pb.query_count() == 51
    pb.run_all()
    lines = pb.report().splitlines()
    stats_at = lines.index("  Monte Carlo statistics over 50 trials")
    assert lines[stats_at + 2:stats_at + 6] == decks.MONTE_TRAN_STATS
    assert lines[stats_at - 3:stats_at - 1] == decks.MONTE_TRAN_TAIL
    # the nominal point: R C ln 2 within the step control's resolution
    m0, r0 = pb.measures(0, 1, ["tc"])
    assert r0.tolist() == [0] and abs(m0[0, 0] - 1e3 * 1e-9 * math.log(2)) < 1e-9


AM_SFFM_SP = """AM and SFFM in ngspice's argument order
VA a 0 AM(0 1 0.5 1k 10k 0.1m 30 45)
VZ z 0 AM(0 0.8 0 1k 10k 0.2m 0 60)
VF f 0 SFFM(0.2 1 2k 3 10k 0.5m 30 10)
RA a 0 1k
RZ z 0 1k
RF f 0 1k
.tran 0.2u 2m
{meas}
.end
"""


def test_am_sffm_cards_follow_ngspice(engine):
    """AM(VO VMO VMA FM FC TD PHASEM PHASEC) read in ngspice's order has
    ngspice's value VO + (VMO + VMA sin(2 pi FM t' + PHASEM)) sin(2 pi FC t'
    + PHASEC), t' = t - TD (vsrcload.c as restated by
    frontend/builder.zig:1547-1562 amArgs); VMA = 0 holds the envelope at
    VMO; SFFM(VO VA FM MDI FC TD PHASEM PHASEC) is VO + VA sin(2 pi FC t' +
    PHASEC + MDI sin(2 pi FM t' + PHASEM)) after TD (models/vsource.va:
    304-317). FIND AT interpolates between accepted points (dt <= 0.2u), so
    the tolerance is the interpolation's, far below any slot mix-up."""
    say("case: AM and SFFM deck cards in ngspice order, FIND AT against the textbook forms")
    at = [0.35e-3, 0.61e-3, 0.9e-3, 1.27e-3, 1.73e-3]
    meas = "\n".join(f".meas tran {nd}{k} FIND v({nd}) AT={t!r}" for nd in "azf" for k, t in enumerate(at))
    pb = E.Problem(AM_SFFM_SP.format(meas=meas))
    names = [f"{nd}{k}" for nd in "azf" for k in range(len(at))]
    pb.run_all()
    m, reasons = pb.measures(0, 1, names)
    assert reasons.tolist() == [0]
    d2r = math.pi / 180
    for k, t in enumerate(at):
        tp = t - 0.1e-3
        a = (1 + 0.5 * math.sin(2 * math.pi * 1e3 * tp + 30 * d2r)) * math.sin(2 * math.pi * 1e4 * tp + 45 * d2r)
        tz = t - 0.2e-3
        z = 0.8 * math.sin(2 * math.pi * 1e4 * tz + 60 * d2r)
        tf = t - 0.5e-3
        f = 0.0 if tf <= 0 else 0.2 + math.sin(2 * math.pi * 1e4 * tf + 10 * d2r
                                               + 3 * math.sin(2 * math.pi * 2e3 * tf + 30 * d2r))
        for nd, want in (("a", a), ("z", z), ("f", f)):
            got = m[0, names.index(f"{nd}{k}")]
            assert abs(got - want) < 2e-3, (nd, t, got, want)


def test_rc_lowpass_ac_plot(engine):
    """rc_lowpass.sp's `.ac dec 10 10 100k` through the points path: the
    AC plot the reference prints (docs/learn/analyses.md:28, 52-56): 41
    points, the columns' names, and its first two rows (frequency, i(v1),
    v(in), v(out), real and imaginary) to a few ulps of each unknown's
    scale, at point 0's operating point."""
    say("case: rc_lowpass.sp .ac dec 10 10 100k, points path, AC print plot")
    pb = E.Problem(decks.RC_LOWPASS_SP)
    pb.run_all()
    text = pb.report(tables=True)
    head, names, rows = parse_print_table(text)
    assert head == "AC Analysis: RC low-pass filter"
    assert names == ["Index", "frequency", "(imag)", "i(v1)", "(imag)", "v(in)", "(imag)", "v(out)", "(imag)"]
    assert rows.shape == (41, 9)
    assert rows[0, 1] == 10.0 and abs(rows[-1, 1] - 1e5) <= 1e-10
    want = [(1e1, -9.979542742416352e-8, -9.989267655685869e-6, 1.0, 0.0, 9.999002045725759e-1,
             -9.98926765568587e-3),
            (1.2589254117941673e1, -1.5815586205860516e-7, -1.2575008897885423e-5, 1.0, 0.0,
             9.998418441379414e-1, -1.2575008897885423e-2)]
    for k, (f, ir, ii, vir, vii, vor, voi) in enumerate(want):
        assert abs(rows[k, 1] - f) <= 2e-15 * f
        assert abs(rows[k, 3] - ir) <= 2e-18 and abs(rows[k, 4] - ii) <= 2e-18
        assert rows[k, 5] == vir and rows[k, 6] == vii
        assert abs(rows[k, 7] - vor) <= 2e-15 and abs(rows[k, 8] - voi) <= 2e-15


def test_rc_lowpass_ac_plot_formats(engine, tmp_path):
    """The AC plot in the reference's other formats (src/output/csv.zig,
    ascii_raw.zig, rawfile.zig): their headers exactly as the reference's
    writers lay them out, and the same numbers as the print plot."""
    say("case: rc_lowpass.sp AC plot as CSV, ASCII raw and binary raw")
    pb = E.Problem(decks.RC_LOWPASS_SP)
    pb.run_all()
    _, _, rows = parse_print_table(pb.report(tables=True))
    vals = rows[:, 1:]
    csv = pb.report(plot_format="csv").splitlines()
    assert csv[0] == ("frequency_re,frequency_im,i(v1)_re,i(v1)_im,v(in)_re,v(in)_im,v(out)_re,v(out)_im")
    # the AC plot's 41 rows (the transient plot follows with its own header)
    assert csv[42] == "time,i(v1),v(in),v(out)"
    assert np.array_equal(np.array([[float(v) for v in ln.split(",")] for ln in csv[1:42]]), vals)
    head = ("Title: RC low-pass filter\nDate: Thu Jan  1 00:00:00 1970\nPlotname: AC Analysis\nFlags: complex\n"
            "No. Variables: 4\nNo. Points: 41\nVariables:\n\t0\tfrequency\tfrequency\n\t1\ti(v1)\tcurrent\n"
            "\t2\tv(in)\tvoltage\n\t3\tv(out)\tvoltage\n")
    asc = pb.report(plot_format="ascii")
    assert asc.startswith(head + "Values:\n")
    lines = asc[len(head) + len("Values:\n"):].splitlines()
    assert lines[41 * 4].startswith("Title: ") and lines[0].startswith("0\t") and lines[4].startswith("1\t")
    first = [complex(*map(float, ln.split("\t")[-1].split(","))) for ln in lines[:4]]
    assert np.array_equal(np.array([c.real for c in first]), vals[0, 0::2])
    b = pb.report(plot_format="binary")
    hb = (head + "Binary:\n").encode()
    assert b[:len(hb)] == hb
    data = np.frombuffer(b[len(hb):len(hb) + 41 * 8 * 8], dtype=np.float64)
    assert np.array_equal(data.reshape(41, 8), vals)
    assert b[len(hb) + 41 * 8 * 8:].startswith(b"Title: RC low-pass filter")


def test_rc_lowpass_tran_plot(engine):
    """rc_lowpass.sp's `.tran 10u 5m` as the transient plot of point 0 (the
    reference's run of it has 519 points, docs/learn/analyses.md:25-30):
    time from 0 to 5 ms, v(in) the PULSE(0 1 0 1u 1u 5m 10m) source at
    every accepted time, and the source's current -(v(in) - v(out)) / 1k."""
    say("case: rc_lowpass.sp transient plot through the points path")
    pb = E.Problem(decks.RC_LOWPASS_SP)
    pb.run_all()
    text = pb.report(tables=True)
    at = text.index("Transient Analysis: RC low-pass filter")
    head, names, rows = parse_print_table(text[at:])
    assert names == ["Index", "time", "i(v1)", "v(in)", "v(out)"]
    assert rows.shape == (519, 5)
    t = rows[:, 1]
    assert t[0] == 0.0 and abs(t[-1] - 5e-3) <= 1e-15 and np.all(np.diff(t) > 0)
    vin = np.clip(t / 1e-6, 0.0, 1.0)
    assert np.allclose(rows[:, 3], vin, rtol=0, atol=1e-12)
    assert np.allclose(rows[:, 2], -(rows[:, 3] - rows[:, 4]) / 1e3, rtol=1e-9, atol=1e-15)


def test_diode_iv_dc_plot(engine):
    """diode_iv.sp's `.dc V1 0 0.8 0.01` through the points path: the DC
    transfer characteristic the reference prints (its title, the sweep
    column v(v-sweep) and the probes, the first three rows), 81 points by
    the reference's sweep count."""
    say("case: diode_iv.sp .dc V1 0 0.8 0.01, points path, DC print plot")
    pb = E.Problem(decks.DIODE_IV_SP)
    pb.run_all()
    text = pb.report(tables=True)
    lines = text.splitlines()
    assert lines[0] == "DC transfer characteristic: Diode I-V curve with a DC sweep"
    assert lines[1] == "Index\tv(v-sweep)\ti(v1)\tv(a)"
    # the sweep and node columns digit for digit; i(v1) to the last bit or
    # two of the diode current (the same bar as the lanes' .dc test: the
    # exponential's last bit differs between the two builds' exp)
    for got, want in zip(lines[3:6], decks.DIODE_IV_ROWS):
        g, w = got.split("\t"), want.split("\t")
        assert [g[0], g[1], g[3]] == [w[0], w[1], w[3]]
        assert abs(float(g[2]) - float(w[2])) <= 4e-16 * abs(float(w[2]))
    _, _, rows = parse_print_table(text)
    assert rows.shape == (81, 4)
    assert abs(rows[-1, 1] - 0.8) <= 1e-15 and np.array_equal(rows[:, 1], rows[:, 3])
    # the forward current grows by e per thermal voltage: monotone, negative
    assert np.all(np.diff(rows[:, 2]) < 0)


@pytest.mark.parametrize("card,count", [
    (".dc V1 0.3 0.95 0.005", 131),  # the ratio lands just under 130 in f64
    (".dc V1 0 1 0.3", 4),           # a step that does not divide the span
    (".dc V1 5 0 -1", 6),            # descending
    (".dc V1 0 1 -0.1", 1),          # a step pointing away from stop
    (".dc I1 0 1m 0.1m", 11),        # a current source: i(i-sweep)
])
def test_dc_sweep_counts(engine, card, count):
    """The reference's sweepCount cases (src/analysis/dc/dc.zig:667-681)
    as decks: the DC plot's number of rows and its sweep column's name."""
    say(f"case: sweep count of {card}")
    src = "I1 0 a 0\n" if card.split()[1] == "I1" else "V1 a 0 0\n"
    pb = E.Problem(f"sweep count\n{src}R1 a 0 1k\n{card}\n.end\n")
    pb.run_all()
    text = pb.report(tables=True)
    lines = text.splitlines()
    assert lines[1].split("\t")[1] == ("i(i-sweep)" if "I1" in card else "v(v-sweep)")
    _, _, rows = parse_print_table(text)
    assert rows.shape[0] == count


def test_ac_only_deck(engine):
    """A deck whose one analysis is `.ac` (rc_lowpass.sp without its `.op`
    and `.tran`): the AC lanes still solve point 0's operating point first,
    and the AC plot equals the full deck's bit for bit."""
    say("case: rc_lowpass.sp with .ac only, points path")
    ac_only = "\n".join(ln for ln in decks.RC_LOWPASS_SP.splitlines() if not ln.startswith((".op", ".tran")))
    full = E.Problem(decks.RC_LOWPASS_SP)
    full.run_all()
    pb = E.Problem(ac_only + "\n")
    pb.run_all()
    h0, n0, r0 = parse_print_table(full.report(tables=True))
    h1, n1, r1 = parse_print_table(pb.report(tables=True))
    assert (h1, n1) == (h0, n0) and r1.shape == (41, 9)
    assert np.array_equal(r1, r0)


def test_cli_dc_rawfile_csv(tmp_path, engine):
    """`espice --rawfile=FILE --format=csv diode_iv.sp`: the DC plot as the
    reference's CSV writer lays it out (src/output/csv.zig): the sweep
    column first, then the probes, 81 rows."""
    say("case: espice --rawfile=FILE --format=csv diode_iv.sp")
    deck = tmp_path / "diode_iv.sp"
    deck.write_text(decks.DIODE_IV_SP)
    raw = tmp_path / "dc.csv"
    exe = os.path.join(ROOT, "egspice_amd", "espice")
    r = subprocess.run([exe, f"--rawfile={raw}", "--format=csv", str(deck)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = raw.read_text().splitlines()
    assert lines[0] == "v(v-sweep),i(v1),v(a)"
    assert len(lines) == 82
    rows = np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])
    assert rows[0].tolist() == [0.0, 0.0, 0.0] and rows[1, 0] == 0.01 and rows[1, 2] == 0.01
    assert abs(rows[1, 1] / -1.386624089968383e-14 - 1) <= 4e-16
