"""PSP103 lanes on the GPU against the oracle, bit for bit: the two-stage
Miller op-amp (8 PSP103 transistors, 34 unknowns) under Monte Carlo inputs,
every lane warm from the nominal operating point, and the gmin ladder
that finds that nominal point from a cold start."""
import numpy as np
import pytest

import egspice_amd as E
import oracle_lib as O
from egspice_amd import workloads as W



def say(msg):
    print(msg, flush=True)

LADDER = [1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10, 1e-11, 1e-12, 0.0]


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    both_nan = np.isnan(a) & np.isnan(b)
    return np.array_equal(a.view(np.uint64)[~both_nan], b.view(np.uint64)[~both_nan])


def engine_opamp(n_lanes, inputs=None):
    t = E.Template()
    inst = W.miller_opamp(t.add, lambda nm: E.param_index(E.PSP103, nm))
    t.freeze()
    lanes = E.Lanes(t, n_lanes)
    if inputs is not None:
        for kind, name, slot, vals in W.opamp_lane_params(inputs, lambda nm: E.param_index(E.PSP103, nm)):
            lanes.set_param(kind, inst[name], slot, vals)
    lanes.prepare()
    return t, lanes


def oracle_opamp(lib, n_lanes, inputs=None):
    c = O.Circuit(lib)
    idx = lambda nm: lib.or_va_param_index(O.PSP103, nm.encode())
    inst = W.miller_opamp(c.add, idx)
    c.freeze()
    L = O.Lanes(c, n_lanes)
    if inputs is not None:
        for kind, name, slot, vals in W.opamp_lane_params(inputs, idx):
            L.set_param(kind, inst[name], slot, vals)
    L.prepare()
    return c, L


@pytest.mark.gpu
def test_opamp_pattern_matches_oracle(engine):
    say("case: op-amp template, engine pattern vs oracle")
    lib = O.load()
    t, _ = engine_opamp(1)
    c, _ = oracle_opamp(lib, 1)
    assert (t.n, t.nnz) == (c.n, c.nnz) == (34, 171)
    assert np.array_equal(t.col_ptr, c.col_ptr) and np.array_equal(t.row_idx, c.row_idx)


@pytest.mark.gpu
def test_opamp_gmin_ladder_bitwise(engine):
    say("case: op-amp nominal operating point by the gmin ladder, 1 lane, GPU vs oracle")
    lib = O.load()
    t, lanes = engine_opamp(1)
    c, L = oracle_opamp(lib, 1)
    xg = np.zeros(t.n)
    xo = np.zeros(c.n)
    for g in LADDER:
        lanes.upload_x(xg.reshape(t.n, 1))
        xg2, sg, ig, _ = lanes.newton(init_fix=False, gmin=g)
        xo2, so, io = L.newton_op(mode=1, x0=xo, init_fix=False, gmin=g)
        assert sg[0] & 1 and so[0] & 1, g
        assert ig[0] == io[0], (g, ig, io)
        assert same_bits(xg2, xo2), g
        xg, xo = xg2[:, 0], xo2[:, 0]


@pytest.mark.gpu
@pytest.mark.parametrize("n_lanes", [64, 1024])
def test_opamp_monte_carlo_warm_bitwise(engine, n_lanes):
    say(f"case: op-amp Monte Carlo, {n_lanes} lanes warm from the nominal point, GPU vs oracle")
    lib = O.load()
    c, L1 = oracle_opamp(lib, 1)
    x = np.zeros(c.n)
    for g in LADDER:
        xs, st, _ = L1.newton_op(mode=1, x0=x, init_fix=False, gmin=g)
        x = xs[:, 0]
    inputs = W.opamp_inputs(n_lanes, seed=11)
    t, lanes = engine_opamp(n_lanes, inputs)
    _, L = oracle_opamp(lib, n_lanes, inputs)
    x0 = np.repeat(x[:, None], n_lanes, axis=1)
    lanes.upload_x(x0)
    xg, sg, ig, ms = lanes.newton(init_fix=False)
    xo, so, io = L.newton_op(mode=1, x0=x0.reshape(-1), init_fix=False, threads=8)
    say(f"  converged {int((sg & 1).sum())}/{n_lanes}, iterations mean {ig.mean():.2f} max {ig.max()}, "
        f"kernel {ms:.3f} ms")
    assert (sg & 1).sum() >= 0.95 * n_lanes
    assert np.array_equal(sg, so) and np.array_equal(ig, io)
    assert same_bits(xg, xo)


@pytest.mark.gpu
def test_opamp_learned_start_bitwise(engine):
    """The learned arm: the warm-start network (trained on engine-solved
    Monte Carlo operating points) starts every lane; the GPU's inference and
    Newton against the oracle's Newton from the same predicted start points,
    bit for bit, and every lane inside the Newton tolerance of its
    nominal-start solution."""
    say("case: op-amp learned warm start, 256 lanes, GPU vs oracle from the network's start points")
    lib = O.load()
    c, L1 = oracle_opamp(lib, 1)
    x = np.zeros(c.n)
    for g in LADDER:
        xs, st, _ = L1.newton_op(mode=1, x0=x, init_fix=False, gmin=g)
        x = xs[:, 0]
    train_in = W.opamp_inputs(1024, seed=5)
    t, tl = engine_opamp(1024, train_in)
    tl.upload_x(np.repeat(x[:, None], 1024, axis=1))
    xt, stt, _, _ = tl.newton(init_fix=False)
    ok = (stt & 1) != 0
    mlp = E.Mlp(8, t.n, hidden=256, depth=3, seed=1)
    mlp.fit_scaling(train_in[ok], xt[:, ok].T)
    mlp.train(train_in[ok], xt[:, ok].T, epochs=40, batch=256, lr=1e-3, seed=2)
    inputs = W.opamp_inputs(256, seed=6)
    _, lanes = engine_opamp(256, inputs)
    xg, sg, ig, _, _ = lanes.newton_learned(mlp, inputs, init_fix=False)
    x0 = mlp.predict(inputs).T.copy()
    _, L = oracle_opamp(lib, 256, inputs)
    xo, so, io = L.newton_op(mode=1, x0=x0.reshape(-1), init_fix=False, threads=8)
    say(f"  converged {int((sg & 1).sum())}/256, iterations mean {ig.mean():.2f} max {ig.max()}")
    assert np.array_equal(sg, so) and np.array_equal(ig, io)
    assert same_bits(xg, xo)
    # against the nominal start: the same operating point within the
    # Newton tolerance (reltol 1e-3 of the node voltage plus vntol)
    lanes.upload_x(np.repeat(x[:, None], 256, axis=1))
    xn, sn, _, _ = lanes.newton(init_fix=False)
    both = ((sg & 1) != 0) & ((sn & 1) != 0)
    nodes = [t.node(nm) for nm in ("vdd", "inp", "nbias", "tail", "n1", "n2", "out")]
    dev = np.abs(xg - xn)[nodes][:, both]
    assert np.all(dev <= 1e-3 * np.abs(xn[nodes][:, both]) + 1e-6)


def opamp_step_template(lib, n_lanes, inputs, engine=True):
    step = (1.0, 10e-9, 1e-9, 500e-9)
    if engine:
        t = E.Template()
        inst = W.miller_opamp(t.add, lambda nm: E.param_index(E.PSP103, nm), step=step)
        t.freeze()
        lanes = E.Lanes(t, n_lanes)
        for kind, name, slot, vals in W.opamp_lane_params(inputs, lambda nm: E.param_index(E.PSP103, nm)):
            lanes.set_param(kind, inst[name], slot, vals)
        lanes.prepare()
        return t, lanes
    c = O.Circuit(lib)
    idx = lambda nm: lib.or_va_param_index(O.PSP103, nm.encode())
    inst = W.miller_opamp(c.add, idx, step=step)
    c.freeze()
    L = O.Lanes(c, n_lanes)
    for kind, name, slot, vals in W.opamp_lane_params(inputs, idx):
        L.set_param(kind, inst[name], slot, vals)
    L.prepare()
    return c, L


@pytest.mark.gpu
def test_opamp_transient_bitwise(engine):
    """A 100 mV input step through the buffer, `.tran 10n 1u`, PSP103 charges
    (gate, bulk, drain and junction charge sites) in the companion and the
    step control: every accepted time and row, and every step statistic,
    bit for bit against the oracle, from the nominal operating point."""
    say("case: op-amp transient .tran 10n 1u, 12 Monte Carlo lanes warm from the nominal point, GPU vs oracle")
    lib = O.load()
    c, L1 = oracle_opamp(lib, 1)
    x = np.zeros(c.n)
    for g in LADDER:
        xs, st, _ = L1.newton_op(mode=1, x0=x, init_fix=False, gmin=g)
        x = xs[:, 0]
    n_lanes = 12
    inputs = W.opamp_inputs(n_lanes, seed=9)
    t, lanes = opamp_step_template(lib, n_lanes, inputs)
    c, L = opamp_step_template(lib, n_lanes, inputs, engine=False)
    lanes.upload_x(np.repeat(x[:, None], n_lanes, axis=1))
    opts = E.tran_opts(1e-6, 10e-9, record_capacity=4096, op_warm=True, max_steps=200_000)
    res, _, times, rows, ms = lanes.tran(opts)
    say(f"  statuses {res['status'].tolist()}, points {res['points'].tolist()}, kernel {ms:.2f} ms")
    oo = O.tran_opts(1e-6, 10e-9, max_steps=200_000)
    out_row = t.node("out")
    for l in range(n_lanes):
        tt, xx, sto = L.tran(l, oo, x0=x)
        k = int(res["points"][l])
        assert res["status"][l] == 0 and sto.status == 0, (l, res["status"][l], sto.status)
        assert sto.points == k and sto.steps == res["steps"][l] and sto.nr_iters == res["nr_iters"][l], l
        assert np.array_equal(times[l, :k].view(np.uint64), tt.view(np.uint64)), f"lane {l} times"
        assert same_bits(rows[l, :k], xx), f"lane {l} rows"
        # the buffer follows its input: out has settled at the pulse's
        # 1.0 V before the pulse ends (511 ns) and is back at 0.9 V at 1 us
        settled = xx[np.searchsorted(tt, 500e-9) - 1, out_row]
        assert abs(settled - 1.0) < 5e-3, settled
        assert abs(xx[-1, out_row] - 0.9) < 5e-3
