"""BSIM4 lanes on the GPU against the oracle, bit for bit: the 31-stage
ring oscillator (62 bsim4va transistors, 34 unknowns) under Monte Carlo
VTH0/TOXE, its cold operating point (the limiter's MODEINITJCT seeds and
limited Newton) and its 5 ns trapezoidal transient with the period's two
crossings as .meas passes."""
import ctypes as C

import numpy as np
import pytest

import egspice_amd as E
import oracle_lib as O
from egspice_amd import workloads as W


def say(msg):
    print(msg, flush=True)


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return np.array_equal(a.view(np.uint64)[~both_nan], b.view(np.uint64)[~both_nan])


def engine_ring(n_lanes, seed=None):
    t = E.Template()
    idx = lambda nm: E.param_index(E.BSIM4, nm)
    inst = W.ring_oscillator(t.add, idx)
    t.freeze()
    lanes = E.Lanes(t, n_lanes)
    if seed is not None:
        for kind, name, slot, vals in W.ring_lane_params(n_lanes, idx, seed=seed):
            lanes.set_param(kind, inst[name], slot, vals)
    lanes.prepare()
    return t, lanes


def oracle_ring(lib, n_lanes, seed=None):
    c = O.Circuit(lib)
    idx = lambda nm: lib.or_va_param_index(O.BSIM4, nm.encode())
    inst = W.ring_oscillator(c.add, idx)
    c.freeze()
    L = O.Lanes(c, n_lanes)
    if seed is not None:
        for kind, name, slot, vals in W.ring_lane_params(n_lanes, idx, seed=seed):
            L.set_param(kind, inst[name], slot, vals)
    L.prepare()
    return c, L


@pytest.mark.gpu
def test_ring_pattern_matches_oracle(engine):
    say("case: ring oscillator template, engine pattern vs oracle")
    lib = O.load()
    t, _ = engine_ring(1)
    c, _ = oracle_ring(lib, 1)
    assert (t.n, t.nnz) == (c.n, c.nnz)
    assert np.array_equal(t.col_ptr, c.col_ptr) and np.array_equal(t.row_idx, c.row_idx)


@pytest.mark.gpu
def test_ring_cold_operating_point_bitwise(engine):
    """From x = 0 with the MODEINITJCT seeds: every lane's status, iteration
    count and solution bit for bit, the limiter active throughout."""
    n_lanes = 256
    say(f"case: ring oscillator cold operating point, {n_lanes} Monte Carlo lanes, GPU vs oracle")
    lib = O.load()
    t, lanes = engine_ring(n_lanes, seed=3)
    c, L = oracle_ring(lib, n_lanes, seed=3)
    lanes.upload_x(np.zeros((t.n, n_lanes)))
    xg, sg, ig, ms = lanes.newton(init_fix=True)
    xo, so, io = L.newton_op(mode=1, x0=np.zeros(c.n * n_lanes), init_fix=True, threads=8)
    say(f"  converged {int((sg & 1).sum())}/{n_lanes}, iterations mean {ig.mean():.2f}, kernel {ms:.3f} ms")
    # a few Monte Carlo lanes need more than the plain rung (the host's
    # ladder after it is outside the lane path); their outcome is compared
    # all the same
    assert (sg & 1).mean() > 0.85
    assert np.array_equal(sg, so) and np.array_equal(ig, io)
    assert same_bits(xg, xo)


@pytest.mark.gpu
def test_ring_transient_bitwise(engine):
    """`.tran 10p 5n` trapezoidal from each lane's cold operating point,
    the kick into n1 at 50 ps: every accepted time and row, every step
    statistic and both crossing passes bit for bit; every lane oscillates,
    a full period between the first two rising crossings of VDD/2 after
    0.2 ns."""
    n_lanes = 8
    say(f"case: ring oscillator transient .tran 10p 5n, {n_lanes} Monte Carlo lanes, GPU vs oracle")
    lib = O.load()
    t, lanes = engine_ring(n_lanes, seed=5)
    c, L = oracle_ring(lib, n_lanes, seed=5)
    n1 = t.node("n1")
    m = W.RING_MEAS
    passes = [E.meas_cross(n1, m["val"], rise=1, td=m["td"]), E.meas_cross(n1, m["val"], rise=2, td=m["td"])]
    opts = E.tran_opts(5e-9, 10e-12, record_capacity=4096, max_steps=200_000)
    res, meas, times, rows, ms = lanes.tran(opts, passes)
    say(f"  statuses {res['status'].tolist()}, points {res['points'].tolist()}, kernel {ms:.2f} ms")
    oo = O.tran_opts(5e-9, 10e-12, max_steps=200_000)
    periods = []
    for l in range(n_lanes):
        tt, xx, sto = L.tran(l, oo)
        k = int(res["points"][l])
        assert res["status"][l] == 0 and sto.status == 0, (l, res["status"][l], sto.status)
        assert sto.points == k and sto.steps == res["steps"][l] and sto.nr_iters == res["nr_iters"][l], l
        assert sto.rej_newton == res["rej_newton"][l] and sto.rej_lte == res["rej_lte"][l], l
        assert np.array_equal(times[l, :k].view(np.uint64), tt.view(np.uint64)), f"lane {l} times"
        assert same_bits(rows[l, :k], xx), f"lane {l} rows"
        for p, rise in ((0, 1), (1, 2)):
            y = np.ascontiguousarray(xx[:, n1])
            want = lib.or_meas_cross(O.f64(tt)[1], O.f64(y)[1], k,
                                     C.byref(O.clause(val=m["val"], rise=rise, td=m["td"])))
            assert same_bits(meas[l, p, 0], want), (l, p)
        period = meas[l, 1, 0] - meas[l, 0, 0]
        periods.append(period)
        assert 1.5e-9 < period < 3e-9, (l, period)
