"""The operating point's ladder after the plain rung (oracle/ladder.c,
src/analysis/dc/op.zig:140-415): its step rules against the reference's
own known answers and recorded descents, the gmin placement against the
engine's (the engine library's template code runs on the host), and
every rung's answer against the circuit's own equations.

Known answers: gminFactor's quarters of the Newton budget
(src/analysis/tests/integration.zig:82-92). Recorded descents:
tests/golden/gmin_ladders.txt is the reference's own fixture
(src/analysis/tests/integration.assets/gmin_ladders.txt): the gmin each
solve of 17 real descents tried, whether Newton converged and in how many
iterations (0 with no convergence: a factor the LU refused). Replaying the
outcomes, the state machine must try the next recorded gmin (printed to 4
digits), and of the descents only chain_bsim4_100k stalls, at solve 29
(integration.zig:94-133)."""
import ctypes as C
import os

import numpy as np
import pytest

import egspice_amd as E
import oracle_lib as O
from egspice_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
NONE = 0xFFFFFFFF


def descents():
    out, name = [], None
    for line in open(os.path.join(HERE, "golden", "gmin_ladders.txt")):
        line = line.strip()
        if line.startswith("== "):
            name = line[3:]
            out.append((name, []))
        elif line.startswith("ladder: gmin="):
            f = line[len("ladder: gmin="):].replace("=", " ").split()
            out[-1][1].append((float(f[0]), f[2] == "true", int(f[4])))
    return out


def test_gmin_factor_known_answers(oracle):
    f = oracle.or_gmin_factor
    assert f(4.0, 25, 100) == min(4 * 4.0 ** 0.5, 10.0)
    assert f(4.0, 26, 100) == 4.0
    assert f(4.0, 75, 100) == 4.0
    assert f(4.0, 76, 100) == 2.0
    assert f(9.0, 1, 100) == 10.0
    assert f(1.00001, 99, 100) == 1.00005


def test_recorded_descents_replay(oracle):
    ds = descents()
    assert len(ds) == 17 and ds[0][0] == "chain_bsim4_100k"
    for name, steps in ds:
        conv = (C.c_int * len(steps))(*[int(s[1]) for s in steps])
        its = np.array([s[2] for s in steps], dtype=np.uint32)
        nxt = np.zeros(len(steps))
        act = (C.c_int * len(steps))()
        k = oracle.or_gmin_replay(conv, its.ctypes.data_as(C.POINTER(C.c_uint32)), len(steps), 1e-2, 100,
                                  nxt.ctypes.data_as(C.POINTER(C.c_double)), act)
        # the first gmin tried is gmin_start / 10
        assert abs(steps[0][0] - 1e-3) <= 5e-4 * 1e-3, name
        for j in range(k):
            if act[j] != 0 or j + 1 >= len(steps):
                break
            want = steps[j + 1][0]
            assert abs(nxt[j] - want) <= 5e-4 * want, (name, j, nxt[j], want)
        # stall detection exactly as the reference's test drives it
        g = np.array([s[0] for s in steps])
        cv = (C.c_int * len(steps))(*[int(s[1]) for s in steps])
        stall = None
        for j in range(len(steps) - 1):
            if oracle.or_gmin_stalled(g.ctypes.data_as(C.POINTER(C.c_double)), cv, j + 1, 1e-2, g[j + 1]):
                stall = j
                break
        assert stall == (29 if name == "chain_bsim4_100k" else None), name
        if name == "chain_bsim4_100k":
            # the replay leaves for source stepping at that same solve
            assert k == 30 and act[29] == 2


def test_stall_needs_twenty_slow_solves(oracle):
    """gminStalled looks only at the last 20 solves: a crawl at factor 1.1
    (under 10^(1/20)) stalls on the 20th, one at 1.13 never does."""
    for factor, want in ((1.1, True), (1.13, False)):
        g = np.array([1e-3 / factor ** k for k in range(20)])
        cv = (C.c_int * 20)(*([1] * 20))
        p = g.ctypes.data_as(C.POINTER(C.c_double))
        assert bool(oracle.or_gmin_stalled(p, cv, 20, 1e-2, g[-1] / factor)) == want
        assert not oracle.or_gmin_stalled(p, cv, 19, 1e-2, g[-2] / factor)


def both_templates(oracle, builder, *args):
    t = E.Template()
    builder(t.add, *args)
    t.freeze()
    c = O.Circuit(oracle)
    builder(c.add, *args)
    c.freeze()
    return t, c


def divider(add):
    add(E.VSOURCE, ["in", "0"], {E.VSOURCE_DC: 10.0})
    add(E.RESISTOR, ["in", "out"], {0: 1e3})
    add(E.RESISTOR, ["out", "0"], {0: 3e3})


def floating(add):
    # a source between two nets that each reach ground through a resistor,
    # then a grounded one on a third net
    add(E.RESISTOR, ["a", "0"], {0: 1e3})
    add(E.RESISTOR, ["b", "0"], {0: 2e3})
    add(E.VSOURCE, ["a", "b"], {E.VSOURCE_DC: 1.0})
    add(E.VSOURCE, ["c", "0"], {E.VSOURCE_DC: 2.0})
    add(E.RESISTOR, ["c", "a"], {0: 5e2})


def test_gmin_stamps_known_answers(oracle):
    """A grounded source's gmin lands on its two incidence entries, not its
    node's diagonal; a floating source, with two twins, swaps with the first
    (its lower row); rows without a source keep their diagonal."""
    t, c = both_templates(oracle, divider)
    col, _ = t.gmin_stamps()
    # rows: in 1, the source's branch 2, out 3
    assert list(col) == [NONE, 2, 1, 3]
    t, c = both_templates(oracle, floating)
    col, _ = t.gmin_stamps()
    # rows: a 1, b 2, branch of V(a,b) 3, c 4, branch of V(c) 5
    assert list(col) == [NONE, 3, 2, 1, 5, 4]


@pytest.mark.parametrize("case", ["divider", "floating", "chain", "diode", "opamp"])
def test_gmin_stamps_engine_matches_oracle(oracle, case):
    """The engine places the stamps structurally (a voltage source's
    incidence); the oracle reads |value| == 1 off lane 0's assembly, as the
    reference does."""
    if case == "opamp":
        idx = lambda nm: oracle.or_va_param_index(O.PSP103, nm.encode())
        t = E.Template()
        W.miller_opamp(t.add, lambda nm: E.param_index(E.PSP103, nm))
        t.freeze()
        c = O.Circuit(oracle)
        W.miller_opamp(c.add, idx)
        c.freeze()
    else:
        b = {"divider": divider, "floating": floating, "chain": W.inverter_chain, "diode": W.diode_rectifier}[case]
        t, c = both_templates(oracle, b)
    L = O.Lanes(c, 1)
    L.prepare()
    col_o, slot_o = L.gmin_stamps()
    col_e, slot_e = t.gmin_stamps()
    assert np.array_equal(col_e, col_o) and np.array_equal(slot_e, slot_o)
    # every stamped slot holds (row, col) in the pattern
    for i in range(t.n):
        if col_e[i] != NONE:
            j = col_e[i]
            assert t.row_idx[slot_e[i]] == i and t.col_ptr[j] <= slot_e[i] < t.col_ptr[j + 1]


def chain_lanes(oracle, vin):
    c = O.Circuit(oracle)
    W.inverter_chain(c.add)
    c.freeze()
    L = O.Lanes(c, len(vin))
    L.set_param(O.VSRC, 1, O.VSRC_DC, vin)
    L.prepare()
    return c, L


def kcl_ok(L, x, lanes):
    """The circuit's equations at each listed lane's answer: every row of
    the residual (KCL currents, source voltages) under 1e-6."""
    for l in lanes:
        rhs, _ = L.eval_lane(int(l), x[:, l])
        assert np.abs(rhs[1:]).max() <= 1e-6, (l, np.abs(rhs[1:]).max())


def test_gmin_stepping_finds_the_plain_answer(oracle):
    """With the plain rung cut to 2 iterations every lane fails it; the gmin
    descent then reaches the deck's gmin and the clean solve lands on the
    same operating point the unhindered plain Newton finds."""
    vin = np.linspace(0.0, 1.8, 8)
    c, L = chain_lanes(oracle, vin)
    xr, sr, _ = L.newton_op(mode=1)
    assert (sr & 1).all()
    c, L = chain_lanes(oracle, vin)
    x, st, it, rung = L.op_ladder(max_iter=2)
    assert (st == 1).all() and (rung == 2).all()
    assert np.abs(x - xr).max() <= 1e-5
    kcl_ok(L, x, range(len(vin)))
    # the plain rung's 2 iterations count, then every solve's
    assert (it > 2).all() and (it < 200).all()


def test_plain_rung_lanes_skip_the_ladder(oracle):
    vin = np.linspace(0.0, 1.8, 4)
    c, L = chain_lanes(oracle, vin)
    xr, sr, ir = L.newton_op(mode=1)
    c, L = chain_lanes(oracle, vin)
    x, st, it, rung = L.op_ladder()
    assert (rung == 1).all() and np.array_equal(it, ir)
    assert np.array_equal(x.view(np.uint64), xr.view(np.uint64))


def test_source_stepping_and_failures(oracle):
    """A 6-iteration stepping budget wedges the gmin descent on some lanes;
    source stepping from the sources-off point then converges where the
    final Newton (6 iterations, warm at full sources) can finish, and the
    rest report no convergence. The converged answers solve the circuit."""
    vin = np.linspace(0.0, 1.8, 4)
    c, L = chain_lanes(oracle, vin)
    x, st, it, rung = L.op_ladder(max_iter=6, ladder_max_iter=6)
    assert list(rung) == [0, 3, 2, 2]
    assert list(st) == [0, 1, 1, 1]
    c2, L2 = chain_lanes(oracle, vin)
    xr, sr, _ = L2.newton_op(mode=1)
    ok = st == 1
    assert np.abs(x[:, ok] - xr[:, ok]).max() <= 1e-5
    kcl_ok(L, x, np.nonzero(ok)[0])


def sweep_values(start, step, npts):
    """The swept values accumulated as the reference steps them
    (dc.zig:373-379: += step, not start + k * step)."""
    v, out = start, []
    for k in range(npts):
        if k:
            v += step
        out.append(v)
    return np.array(out)


def test_dc_sweep_lanes_diode_golden(oracle):
    """docs/examples/diode_iv.sp's printed rows (docs/learn/analyses.md:
    88-93) on every lane of a `.dc` sweep over lanes: i(v1) at 10 and 20 mV
    to 4e-16."""
    c = O.Circuit(oracle)
    c.add(O.VSRC, ["a", "0"], {O.VSRC_DC: 0.0})
    c.add(O.DIODE, ["a", "0"], {O.DIO["IS"]: 1e-14, O.DIO["N"]: 1.0})
    c.freeze()
    L = O.Lanes(c, 3)
    L.prepare()
    x, conv = L.dc_sweep_lanes(O.VSRC, 0, sweep_values(0.0, 0.01, 81))
    assert conv.all()
    golden = {1: -1.386624089968383e-14, 2: -3.166792969038269e-14}
    for k, v in golden.items():
        assert (np.abs(x[k, 2] / v - 1) <= 4e-16).all()
        assert (np.abs(x[k, 1] - 0.01 * k) <= 1e-17).all()


def test_dc_sweep_lanes_chain(oracle):
    """The inverter chain swept 0-1.8 V on 6 Monte Carlo lanes (VTO spread):
    every point converges, each lane's curve is its own lane swept alone,
    and each point equals that point's cold operating point within the
    Newton tolerance (reltol 1e-3 of the node voltage, plus 1e-5);
    with the warm budget cut to 2 iterations the warm solves fail and the
    cold ladder still finds every point."""
    vin = sweep_values(0.0, 0.1, 19)
    vto = np.array([0.65, 0.68, 0.7, 0.72, 0.75, 0.7])

    def lanes(n_lanes, vt):
        c = O.Circuit(oracle)
        W.inverter_chain(c.add)
        c.freeze()
        L = O.Lanes(c, n_lanes)
        for i in range(0, 10, 2):
            L.set_param(O.MOS1, i, O.M1["VTO"], vt)
        L.prepare()
        return c, L

    c, L = lanes(6, vto)
    x, conv = L.dc_sweep_lanes(O.VSRC, 1, vin)
    assert conv.all()
    c1, L1 = lanes(1, vto[3:4])
    x1, conv1 = L1.dc_sweep_lanes(O.VSRC, 1, vin)
    assert np.array_equal(x[:, :, 3].view(np.uint64), x1[:, :, 0].view(np.uint64))
    for k in (0, 7, 9, 18):
        cc, Lc = lanes(6, vto)
        Lc.set_param(O.VSRC, 1, O.VSRC_DC, np.full(6, vin[k]))
        Lc.prepare()
        xc, sc, _ = Lc.newton_op(mode=1)
        assert (sc & 1).all()
        assert within_tol(x[k], xc), k
    c, L = lanes(6, vto)
    x2, conv2 = L.dc_sweep_lanes(O.VSRC, 1, vin, ladder_max_iter=2)
    assert conv2.all()
    assert within_tol(x2, x)


def within_tol(a, b):
    """Two Newton answers of the same equations from different starts:
    within reltol (1e-3) of the larger magnitude, plus 1e-5."""
    return bool(np.all(np.abs(a - b) <= 1e-3 * np.maximum(np.abs(a), np.abs(b)) + 1e-5))
