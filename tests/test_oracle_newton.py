"""The oracle's Newton drivers and device bodies, pinned by the reference's
golden outputs and by mathematics:
  - docs/examples/diode_iv.sp DC rows as the reference printed them
    (docs/learn/analyses.md:88-93), to a few ulps;
  - the Jacobian of every device body against central finite differences
    (the reference's device/tests/jacobian_fd.zig property);
  - KCL at every converged lane: an independent numpy evaluation of the
    inverter chain's branch currents balances at the published x;
  - the lane-tape driver (the reference's GPU lane algorithm) agrees with
    the CPU path within the Newton tolerance.
"""
import numpy as np
import pytest

import oracle_lib as O
from egspice_amd import workloads as W


def diode_deck(lib):
    c = O.Circuit(lib)
    c.add(O.VSRC, ["a", "0"], {O.VSRC_DC: 0.0})
    c.add(O.DIODE, ["a", "0"], {O.DIO["IS"]: 1e-14, O.DIO["N"]: 1.0})
    return c.freeze()


def test_diode_iv_golden_rows(oracle):
    c = diode_deck(oracle)
    L = O.Lanes(c, 1)
    L.prepare()
    xs, conv = L.dc_sweep(0, 0.0, 0.01, 81)
    assert conv.all()
    br = 2  # ground 0, a 1, the V1 branch 2
    assert xs[0][br] == 0.0 and xs[0][1] == 0.0
    golden = {1: -1.386624089968383e-14, 2: -3.166792969038269e-14}
    for k, v in golden.items():
        assert abs(xs[k][1] - 0.01 * k) <= 1e-17
        assert abs(xs[k][br] / v - 1) <= 4e-16, (k, xs[k][br], v)


def fd_jacobian(L, c, x, h=1e-7):
    """Central differences of the residual, column by column."""
    n = c.n
    J = np.zeros((n, n))
    for j in range(n):
        xp, xm = x.copy(), x.copy()
        step = h * max(1.0, abs(x[j]))
        xp[j] += step
        xm[j] -= step
        fp, _ = L.eval_lane(0, xp)
        fm, _ = L.eval_lane(0, xm)
        J[:, j] = (fp - fm) / (2 * step)
    return J


def csc_dense(c, vals):
    A = np.zeros((c.n, c.n))
    for j in range(c.n):
        for p in range(c.col_ptr[j], c.col_ptr[j + 1]):
            A[c.row_idx[p], j] += vals[p]
    return A


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_jacobian_is_the_finite_difference(oracle, seed):
    c = O.Circuit(oracle)
    W.inverter_chain(c.add)
    c.freeze()
    rng = np.random.default_rng(seed)
    L = O.Lanes(c, 1)
    L.prepare()
    nodes = [c.node(k) for k in ["vdd", "in", "n1", "n2", "n3", "n4", "n5"]]
    for _ in range(20):
        # node voltages inside the rails keep every junction reverse biased;
        # the branch currents are free
        x = rng.uniform(-1e-3, 1e-3, c.n)
        x[0] = 0.0
        x[nodes] = rng.uniform(0.0, 1.8, len(nodes))
        x[c.node("vdd")] = 1.8
        f, vals = L.eval_lane(0, x)
        A = csc_dense(c, vals)
        J = fd_jacobian(L, c, x)
        # column 0 is the ground pin: device entries there land in the
        # trash slot by construction, only the ground stamp's +1 is kept
        A, J = A[:, 1:], J[:, 1:]
        scale = np.abs(A).max()
        assert np.allclose(A, J, rtol=1e-5, atol=1e-7 * scale)


def mos_current(type_, kp, vto, w, l, vd, vg, vs):
    """Shichman-Hodges drain current (mos1.va:338-345, gamma = lambda = 0),
    written independently of the oracle."""
    beta = kp * w / l
    vgs, vds = type_ * (vg - vs), type_ * (vd - vs)
    mode = 1.0
    if vds < 0:
        mode, vgs, vds = -1.0, vgs - vds, -vds
    vgst = vgs - type_ * vto
    if vgst <= 0:
        i = 0.0
    elif vgst <= vds:
        i = 0.5 * beta * vgst * vgst
    else:
        i = beta * vds * (vgst - 0.5 * vds)
    return type_ * mode * i


def test_converged_lanes_satisfy_kcl(oracle):
    c = O.Circuit(oracle)
    W.inverter_chain(c.add)
    c.freeze()
    vin = W.dc_sweep_points(64)
    L = O.Lanes(c, 64)
    L.set_param(O.VSRC, 1, O.VSRC_DC, vin)
    L.prepare()
    x, st, it = L.newton_op(mode=1)
    assert (st & 1).all()
    node = {name: c.node(name) for name in ["vdd", "in", "n1", "n2", "n3", "n4", "n5"]}
    for l in range(64):
        v = {k: x[r, l] for k, r in node.items()}
        assert abs(v["vdd"] - 1.8) < 1e-12 and abs(v["in"] - vin[l]) < 1e-12
        prev = "in"
        for k in range(1, 6):
            out = f"n{k}"
            i_n = mos_current(1.0, 1e-4, 0.7, 2e-6, 1e-6, v[out], v[prev], 0.0)
            i_p = mos_current(-1.0, 4e-5, -0.7, 4e-6, 1e-6, v[out], v[prev], v["vdd"])
            # junction leakage is below 1e-13 A here; the channel currents
            # balance to the Newton tolerance
            assert abs(i_n + i_p) <= 1e-9 + 1e-3 * max(abs(i_n), abs(i_p)), (l, k, i_n, i_p)
            prev = out


def test_lane_tape_driver_agrees_with_cpu_path(oracle):
    c = O.Circuit(oracle)
    W.inverter_chain(c.add)
    c.freeze()
    vin = W.dc_sweep_points(64)
    L = O.Lanes(c, 64)
    L.set_param(O.VSRC, 1, O.VSRC_DC, vin)
    L.prepare()
    x0, s0, _ = L.newton_op(mode=0)
    x1, s1, _ = L.newton_op(mode=1)
    assert (s0 & 1).all() and (s1 & 1).all()
    # both publish x_k inside the Newton tolerance of the same root
    assert np.abs(x0 - x1).max() <= 1e-5


def test_threaded_driver_is_the_serial_one(oracle):
    """Lanes are independent: dealing them over threads changes no bit."""
    c = O.Circuit(oracle)
    W.inverter_chain(c.add)
    c.freeze()
    rng = np.random.default_rng(4)
    N = 96
    L = O.Lanes(c, N)
    L.set_param(O.VSRC, 1, O.VSRC_DC, rng.uniform(0, 1.8, N))
    L.set_param(O.MOS1, 1, O.M1["VTO"], -0.7 + 0.02 * rng.standard_normal(N))
    L.prepare()
    ref = L.newton_op(mode=1)
    for threads in (2, 7):
        got = L.newton_op(mode=1, threads=threads)
        assert np.array_equal(ref[0].view(np.uint64), got[0].view(np.uint64))
        assert np.array_equal(ref[1], got[1]) and np.array_equal(ref[2], got[2])
