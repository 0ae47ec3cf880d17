"""The engine's own per-lane code (egspice_amd/csrc), compiled for the host,
against the oracle's CPU path: the template pattern and column order, and
the Newton outcome of every lane bit for bit, for both the thread-per-lane
body and a host emulation of the workgroup-per-lane wave phases. This pins
the device code's arithmetic on the CPU; the GPU tests pin the same code on
the MI355X."""
import numpy as np
import pytest

import hostlane_lib as H
import oracle_lib as O
from egspice_amd import workloads as W


def mixed(add):
    """A divider into a diode with a current sink and a capacitor: every
    kind but the MOSFET, and a pattern that is not tridiagonal."""
    add(O.VSRC, ["a", 0], {O.VSRC_DC: 5.0})
    add(O.RES, ["a", "b"], {0: 1e3})
    add(O.DIODE, ["b", "c"], {O.DIO["IS"]: 1e-14, O.DIO["N"]: 1.5})
    add(O.RES, ["c", 0], {0: 220.0})
    add(O.RES, ["b", "c"], {0: 1e5})
    add(O.ISRC, ["b", 0], {0: 1e-3})
    add(O.CAP, ["b", 0], {0: 1e-12})
    add(O.DIODE, ["c", 0], {O.DIO["IS"]: 2e-14, O.DIO["N"]: 1.0})
    return 0


def diode_only(add):
    add(O.VSRC, ["a", 0], {O.VSRC_DC: 0.0})
    add(O.DIODE, ["a", 0], {O.DIO["IS"]: 1e-14, O.DIO["N"]: 1.0})
    return 0


def both(oracle, hostlane, builder, *args):
    c = O.Circuit(oracle)
    builder(c.add, *args)
    c.freeze()
    t = H.HostTemplate(hostlane)
    builder(t.add, *args)
    t.freeze()
    return c, t


def bits(a):
    return np.asarray(a, dtype=np.float64).view(np.uint64)


def run_all(c, t, N, params):
    ol, hl = O.Lanes(c, N), H.HostLanes(t, N)
    for L in (ol, hl):
        for kind, inst, k, vals in params:
            L.set_param(kind, inst, k, vals)
        L.prepare()
    xo, so, io = ol.newton_op(mode=1)
    xt, st, it = hl.newton_op()
    xw, sw, iw = hl.newton_op(wave=True)
    return (xo, so, io), (xt, st, it), (xw, sw, iw)


def assert_same(ref, got):
    assert np.array_equal(ref[1], got[1])
    assert np.array_equal(ref[2], got[2])
    assert np.array_equal(bits(ref[0]), bits(got[0]))


@pytest.mark.parametrize("stages", [1, 5, 12])
def test_inverter_chain_pattern_and_lanes(oracle, hostlane, stages):
    c, t = both(oracle, hostlane, W.inverter_chain, stages)
    assert np.array_equal(c.col_ptr, t.col_ptr) and np.array_equal(c.row_idx, t.row_idx)
    assert np.array_equal(c.diag, t.diag)
    vin = W.dc_sweep_points(64)
    ref, thr, wav = run_all(c, t, 64, [(O.VSRC, 1, O.VSRC_DC, vin)])
    assert (ref[1] & 1).all()
    assert_same(ref, thr)
    assert_same(ref, wav)


def test_monte_carlo_lanes(oracle, hostlane):
    c, t = both(oracle, hostlane, W.inverter_chain)
    rng = np.random.default_rng(3)
    N = 200
    params = [(O.VSRC, 1, O.VSRC_DC, rng.uniform(0, 1.8, N)),
              (O.MOS1, 0, O.M1["VTO"], 0.7 + 0.02 * rng.standard_normal(N)),
              (O.MOS1, 3, O.M1["KP"], 4e-5 * (1 + 0.05 * rng.standard_normal(N))),
              (O.MOS1, 2, O.M1["TOX"], 1e-7 * (1 + 0.02 * rng.standard_normal(N))),
              (O.MOS1, 4, O.M1["LAMBDA"], rng.uniform(0, 0.05, N))]
    ref, thr, wav = run_all(c, t, N, params)
    assert (ref[1] & 1).mean() > 0.95
    assert_same(ref, thr)
    assert_same(ref, wav)


def test_tridiagonal_thomas_path(oracle, hostlane):
    c, t = both(oracle, hostlane, diode_only)
    ref, thr, wav = run_all(c, t, 81, [(O.VSRC, 0, O.VSRC_DC, np.arange(81) * 0.01)])
    assert (ref[1] & 1).all()
    assert_same(ref, thr)
    assert_same(ref, wav)


def test_mixed_kinds(oracle, hostlane):
    c, t = both(oracle, hostlane, mixed)
    rng = np.random.default_rng(9)
    N = 48
    params = [(O.VSRC, 0, O.VSRC_DC, rng.uniform(-2, 8, N)),
              (O.ISRC, 0, 0, rng.uniform(-2e-3, 2e-3, N)),
              (O.DIODE, 1, O.DIO["IS"], 10 ** rng.uniform(-16, -12, N))]
    ref, thr, wav = run_all(c, t, N, params)
    assert (ref[1] & 1).all()
    assert_same(ref, thr)
    assert_same(ref, wav)


def test_max_iter_and_no_init_fix(oracle, hostlane):
    """Lanes stopped by the iteration cap report the same status, count and
    last iterate as the oracle, with and without the INITFIX phase."""
    c, t = both(oracle, hostlane, W.inverter_chain)
    vin = W.dc_sweep_points(16)
    for kw in ({"max_iter": 3}, {"max_iter": 7, "init_fix": False}, {"init_fix": False}):
        ol, hl = O.Lanes(c, 16), H.HostLanes(t, 16)
        for L in (ol, hl):
            L.set_param(O.VSRC, 1, O.VSRC_DC, vin)
            L.prepare()
        ref = ol.newton_op(mode=1, **kw)
        assert_same(ref, hl.newton_op(**kw))
        assert_same(ref, hl.newton_op(wave=True, **kw))
