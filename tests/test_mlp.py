"""The learned warm start (include/espice_mlp.h, egspice_amd/warm.py).

CPU: the numpy restatement (oracle/mlp.py) is pinned by central finite
differences of its FP64 form and its bfloat16 rounding by known answers.

GPU: the network's forward pass and its gradients against that
restatement with the same BF16 roundings (FP32 sums on the GPU, FP64 in the
restatement: agreement to a stated tolerance, not bitwise), training
reduces the loss, and the warm start on the 5-stage inverter chain (the
64-point DC sweep) converges every lane in fewer Newton ticks than the cold
start, to the same operating points within the Newton tolerance.
"""
import os
import sys

import numpy as np
import pytest

import egspice_amd as E
from egspice_amd import warm
from egspice_amd import workloads as W

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp as M  # noqa: E402


def say(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ CPU


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # bf16 spacing at 1
    cases = [(1.0, 1.0), (1.0 + 2 ** -8, 1.0), (1.0 + 3 * 2 ** -8, 1.0 + 2 ** -6), (1.0 + 2 ** -8 + 2 ** -20, 1.0 + 2 ** -7),
             (-1.5, -1.5), (3.140625, 3.140625), (0.0, 0.0)]
    for v, want in cases:
        assert M.bf16(np.float32(v)) == np.float32(want), v
    assert M.bf16(one + ulp) == one + ulp


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, hidden, depth, n_out = 3, 7, 2, 4
    sizes = M.shapes(n_in, hidden, depth, n_out)
    params = rng.normal(0, 0.5, sum(fo * fi + fo for fo, fi in sizes))
    x = rng.normal(size=(9, n_in))
    y = rng.normal(size=(9, n_out))
    im, isd = x.mean(0), x.std(0)
    om, osd = y.mean(0), y.std(0)

    def loss(p):
        return M.loss_and_grads(M.unpack(p, n_in, hidden, depth, n_out), x, y, im, isd, om, osd, bf16_ops=False)[0]

    _, g = M.loss_and_grads(M.unpack(params, n_in, hidden, depth, n_out), x, y, im, isd, om, osd, bf16_ops=False)
    h = 1e-6
    fd = np.array([(loss(params + h * e) - loss(params - h * e)) / (2 * h) for e in np.eye(params.size)])
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-9)


def test_oracle_constant_columns():
    """A zero std input contributes 0; a zero std output is exactly its mean."""
    layers = M.unpack(np.ones(sum(fo * fi + fo for fo, fi in M.shapes(2, 3, 1, 2))), 2, 3, 1, 2)
    y = M.predict(layers, np.array([[5.0, 1.0]]), [5.0, 0.0], [0.0, 1.0], [0.25, 2.0], [0.0, 1.0])
    assert y[0, 0] == 0.25


def test_library_exports_the_network():
    lib = E.library()
    for name in ("espice_mlp_create", "espice_mlp_train", "espice_mlp_gradients", "espice_mlp_predict",
                 "espice_lanes_warm_start"):
        assert hasattr(lib, name)


# ------------------------------------------------------------------ GPU


def scaled_problem(seed, n, n_in=8, n_out=12):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 2.0, (n, n_in))
    A = rng.normal(size=(n_in, n_out))
    y = np.tanh(x @ A * 0.5) + 0.1 * x[:, :1] ** 2
    return x, y


@pytest.mark.gpu
def test_mlp_forward_matches_oracle(engine):
    say("case: MLP 8->3x256->12 forward, 300 rows, vs the BF16 restatement")
    x, y = scaled_problem(1, 300)
    net = E.Mlp(8, 12, hidden=256, depth=3, seed=3)
    net.fit_scaling(x, y)
    got = net.predict(x)
    layers = M.unpack(net.params(), 8, 256, 3, 12)
    want = M.predict(layers, x, x.mean(0), x.std(0), y.mean(0), y.std(0))
    scale = y.std(0)
    err = np.abs(got - want) / scale
    say(f"  max |gpu - oracle| / std(y) = {err.max():.2e}")
    # FP32 MFMA sums against the restatement's FP64 sums of the same BF16
    # operands: where the two sums straddle a BF16 rounding boundary an
    # activation lands one BF16 step (2^-8 relative) away; measured 1.0e-3
    # of std(y) at most over 300 rows and 12 outputs
    assert err.max() < 2e-3
    assert np.all(np.isfinite(got))


@pytest.mark.gpu
def test_mlp_gradients_match_oracle(engine):
    say("case: MLP 8->3x256->12 gradients over 200 rows vs the BF16 restatement")
    x, y = scaled_problem(2, 200)
    net = E.Mlp(8, 12, hidden=256, depth=3, seed=5)
    net.fit_scaling(x, y)
    g, loss = net.gradients(x, y)
    layers = M.unpack(net.params(), 8, 256, 3, 12)
    want_loss, want = M.loss_and_grads(layers, x, y, x.mean(0), x.std(0), y.mean(0), y.std(0))
    assert loss == pytest.approx(want_loss, rel=1e-3)
    at = 0
    for k, (fo, fi) in enumerate(M.shapes(8, 256, 3, 12)):
        for part, size in (("W", fo * fi), ("b", fo)):
            a, b = g[at:at + size], want[at:at + size]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            say(f"  layer {k} {part}: relative error {rel:.2e}")
            # FP32 MFMA sums against FP64 ones over BF16 operands: measured
            # 3.9e-3 at most (layer 0's bias, a sum over every row of the batch)
            assert rel < 8e-3, (k, part, rel)
            at += size
    assert at == g.size


@pytest.mark.gpu
def test_mlp_training_reduces_loss(engine):
    say("case: MLP training on 4096 rows, 60 epochs")
    x, y = scaled_problem(3, 4096)
    net = E.Mlp(8, 12, hidden=256, depth=3, seed=7)
    net.fit_scaling(x, y)
    loss = net.train(x, y, epochs=60, batch=256, lr=1e-3)
    say(f"  loss {loss[0]:.3e} -> {loss[-1]:.3e}")
    assert np.all(np.isfinite(loss))
    assert loss[-1] < 0.05 * loss[0]
    xt, yt = scaled_problem(3, 4096)
    pred = net.predict(xt[:512])
    assert np.sqrt(np.mean(((pred - yt[:512]) / yt.std(0)) ** 2)) < 0.2


def chain(N):
    t = E.Template()
    W.inverter_chain(t.add)
    t.freeze()
    return t


def set_vin(lanes, params):
    lanes.set_param(E.VSOURCE, 1, E.VSOURCE_DC, np.asarray(params, dtype=np.float64)[:, 0])


@pytest.mark.gpu
def test_learned_warm_start_inverter_chain(engine):
    say("case: learned warm start on the 64-point inverter-chain sweep (trained on 2048 cold OPs)")
    t = chain(64)
    rng = np.random.default_rng(11)
    train = rng.uniform(0.0, 1.8, (2048, 1))
    net, info = warm.train_warm_start(t, set_vin, train, batch=256)
    say(f"  training: {info}")
    assert info["loss_last"] < info["loss_first"]
    sweep = W.dc_sweep_points(64).reshape(-1, 1)
    cold = E.Lanes(t, 64)
    set_vin(cold, sweep)
    cold.prepare()
    xc, stc, itc, _ = cold.newton_op()
    assert (stc & 1).all()
    lanes = E.Lanes(t, 64)
    set_vin(lanes, sweep)
    lanes.prepare()
    x, st, it, ms = warm.learned_op(lanes, net, sweep)
    say(f"  Newton ticks: cold sum {itc.sum()} max {itc.max()}; learned sum {it.sum()} max {it.max()}; {ms:.3f} ms")
    assert (st & 1).all()
    assert it.sum() < 0.6 * itc.sum() and it.max() < itc.max()
    # the same operating points within the Newton tolerance (reltol 1e-3,
    # vntol 1e-6 on node rows)
    diff = np.abs(x - xc)
    bound = 1e-3 * np.maximum(np.abs(x), np.abs(xc)) + 1e-6
    rows = [t.node(nm) for nm in ["in", "vdd"] + [f"n{k}" for k in range(1, 6)]]
    say(f"  max |learned - cold| on node rows: {diff[rows].max():.3e} V")
    assert np.all(diff[rows] <= bound[rows])


@pytest.mark.gpu
def test_newton_learned_equals_its_two_calls(engine):
    """espice_lanes_newton_learned is espice_lanes_warm_start followed by
    espice_lanes_newton_op with init_fix 0: the same lanes bit for bit."""
    say("case: newton_learned vs warm_start + newton_op, 64-lane inverter chain")
    t = chain(64)
    rng = np.random.default_rng(5)
    net, _ = warm.train_warm_start(t, set_vin, rng.uniform(0.0, 1.8, (512, 1)), batch=256,
                                   schedule=((200, 1e-3),))
    sweep = W.dc_sweep_points(64).reshape(-1, 1)
    a = E.Lanes(t, 64)
    set_vin(a, sweep)
    a.prepare()
    a.warm_start(net, sweep)
    xa, sa, ia, _ = a.newton(init_fix=False)
    b = E.Lanes(t, 64)
    set_vin(b, sweep)
    b.prepare()
    xb, sb, ib, wms, nms = b.newton_learned(net, sweep, init_fix=False)
    assert wms > 0 and nms > 0
    assert np.array_equal(sa, sb) and np.array_equal(ia, ib)
    assert np.array_equal(xa.view(np.uint64), xb.view(np.uint64))


@pytest.mark.gpu
def test_fused_warm_start_equals_layer_by_layer(engine):
    """The warm start's one-launch network (mlp_infer_kernel) writes into
    the lanes exactly what espice_mlp_predict's layer-by-layer kernels
    return: the same MFMA sequences, bias, tanh and roundings, bit for bit,
    over 100 rows (a partial last tile of 32)."""
    say("case: fused warm-start inference vs layer-by-layer predict, 100 rows, 1->3x256->10")
    t = chain(64)
    rng = np.random.default_rng(8)
    net = E.Mlp(1, t.n, hidden=256, depth=3, seed=9)
    xs = rng.uniform(0.0, 1.8, (300, 1))
    net.fit_scaling(xs, rng.normal(size=(300, t.n)))
    params = rng.uniform(0.0, 1.8, (100, 1))
    lanes = E.Lanes(t, 100)
    set_vin(lanes, params)
    lanes.prepare()
    lanes.warm_start(net, params)
    got = lanes.download_x()
    want = net.predict(params).T
    assert np.array_equal(got.view(np.uint64), np.ascontiguousarray(want).view(np.uint64))
