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
