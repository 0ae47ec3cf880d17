"""The lanes' static Newton with FP32 factors and FP64 iterative
refinement (espice_lanes_newton_mixed) against the FP64 oracle: per lane
the error of the operating point, the FP32 solves each tick took, and the
ticks handed back to the FP64 factor. The refinement stops when the
residual meets the FP64 solve's own backward-error bar (LAPACK dsgesv's
test), so on a linear circuit (one Newton step) the answer is the FP64
one to within the conditioning times that bar; on the nonlinear circuits
the converged lanes lie within the Newton tolerance of the oracle's
(reltol 1e-3 of the node voltage plus vntol 1e-6), measured and printed
per case."""
import numpy as np
import pytest

import egspice_amd as E
import oracle_lib as O
from egspice_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def within_newton_tol(a, b):
    return bool(np.all(np.abs(a - b) <= 1e-3 * np.maximum(np.abs(a), np.abs(b)) + 1e-6))


def spread_ladder(add, stages=12):
    """A resistor ladder whose series resistors alternate 1 ohm and 1 Mohm
    over 1 kohm shunts, driven by 1 V: a conditioning of about 1e6, where
    an FP32 solve alone is good to a few digits."""
    add(E.VSOURCE, ["n0", "0"], {E.VSOURCE_DC: 1.0})
    for k in range(stages):
        add(E.RESISTOR, [f"n{k}", f"n{k + 1}"], {0: 1.0 if k % 2 == 0 else 1e6})
        add(E.RESISTOR, [f"n{k + 1}", "0"], {0: 1e3})


def test_mixed_linear_ladder_is_fp64(oracle, engine):
    N = 64
    say(f"case: mixed Newton, resistor ladder 1 ohm / 1 Mohm, {N} lanes")
    t = E.Template()
    spread_ladder(t.add)
    t.freeze()
    g = E.Lanes(t, N)
    scale = np.linspace(0.5, 2.0, N)
    g.set_param(E.VSOURCE, 0, E.VSOURCE_DC, scale)
    g.prepare()
    c = O.Circuit(oracle)
    spread_ladder(c.add)
    c.freeze()
    o = O.Lanes(c, N)
    o.set_param(O.VSRC, 0, O.VSRC_DC, scale)
    o.prepare()
    g.upload_x(np.zeros((t.n, N)))
    xm, sm, im, solves, fallbacks, ms = g.newton_mixed(init_fix=False)
    xo, so, io = o.newton_op(mode=1, init_fix=False)
    rel = np.abs(xm - xo) / np.maximum(np.abs(xo), 1e-300)
    node = np.abs(xo) > 1e-9
    say(f"  converged {int((sm & 1).sum())}/{N}, FP32 solves per tick {solves.sum() / im.sum():.2f}, "
        f"fallbacks {int(fallbacks.sum())}, max relative error {rel[node].max():.2e}, kernel {ms:.3f} ms")
    assert (sm == 1).all() and (so & 1).all()
    assert np.array_equal(im, io)
    assert fallbacks.sum() == 0
    # refined to FP64: the backward-error bar (sqrt(n) 2^-53) times the
    # ladder's conditioning (about 1e6) bounds the difference from the FP64
    # factor's answer
    assert rel[node].max() <= 1e-9
    # and the refinement did work: more than one FP32 solve per tick
    assert solves.sum() > im.sum()


def test_mixed_inverter_chain_monte_carlo(oracle, engine):
    N = 256
    say(f"case: mixed Newton, inverter chain Monte Carlo, {N} lanes from a cold start")
    t = E.Template()
    W.inverter_chain(t.add)
    t.freeze()
    c = O.Circuit(oracle)
    W.inverter_chain(c.add)
    c.freeze()
    rng = np.random.default_rng(7)
    vin = rng.uniform(0.0, 1.8, N)
    vto = rng.normal(0.7, 0.03, N)
    g = E.Lanes(t, N)
    o = O.Lanes(c, N)
    g.set_param(E.VSOURCE, 1, E.VSOURCE_DC, vin)
    o.set_param(O.VSRC, 1, O.VSRC_DC, vin)
    for i in range(0, 10, 2):
        g.set_param(E.MOS1, i, E.MOS1_PARAMS["VTO"], vto)
        o.set_param(O.MOS1, i, O.M1["VTO"], vto)
    g.prepare()
    o.prepare()
    g.upload_x(np.zeros((t.n, N)))
    xm, sm, im, solves, fallbacks, ms = g.newton_mixed()
    xo, so, io = o.newton_op(mode=1)
    err = np.abs(xm - xo).max(axis=0)
    say(f"  converged {int((sm & 1).sum())}/{N} (FP64 {int((so & 1).sum())}), same ticks {int((im == io).sum())}, "
        f"FP32 solves per tick {solves.sum() / max(im.sum(), 1):.2f}, fallbacks {int(fallbacks.sum())}, "
        f"max node error {err.max():.2e} V, median {np.median(err):.2e} V, kernel {ms:.3f} ms")
    both = ((sm & 1) != 0) & ((so & 1) != 0)
    assert both.sum() >= 0.99 * N
    assert within_newton_tol(xm[:, both], xo[:, both])


def test_mixed_opamp_monte_carlo(oracle, engine):
    N = 128
    say(f"case: mixed Newton, psp103 op-amp Monte Carlo, {N} lanes warm from the nominal point")
    inputs = W.opamp_inputs(N, seed=13)
    t = E.Template()
    eidx = lambda nm: E.param_index(E.PSP103, nm)
    inst = W.miller_opamp(t.add, eidx)
    t.freeze()
    g1 = E.Lanes(t, 1)
    g1.prepare()
    xn, sn, _, _, _ = g1.op_ladder()
    assert sn[0] == 1
    g = E.Lanes(t, N)
    for kind, name, slot, vals in W.opamp_lane_params(inputs, eidx):
        g.set_param(kind, inst[name], slot, vals)
    g.prepare()
    oidx = lambda nm: oracle.or_va_param_index(O.PSP103, nm.encode())
    c = O.Circuit(oracle)
    oinst = W.miller_opamp(c.add, oidx)
    c.freeze()
    o = O.Lanes(c, N)
    for kind, name, slot, vals in W.opamp_lane_params(inputs, oidx):
        o.set_param(kind, oinst[name], slot, vals)
    o.prepare()
    x0 = np.repeat(xn, N, axis=1)
    g.upload_x(x0)
    xm, sm, im, solves, fallbacks, ms = g.newton_mixed(init_fix=False)
    xo, so, io = o.newton_op(mode=1, x0=x0.reshape(-1), init_fix=False, threads=8)
    err = np.abs(xm - xo).max(axis=0)
    say(f"  converged {int((sm & 1).sum())}/{N}, same ticks {int((im == io).sum())}, "
        f"FP32 solves per tick {solves.sum() / max(im.sum(), 1):.2f}, fallbacks {int(fallbacks.sum())}, "
        f"max node error {err.max():.2e} V, median {np.median(err):.2e} V, kernel {ms:.3f} ms")
    both = ((sm & 1) != 0) & ((so & 1) != 0)
    assert both.sum() >= 0.95 * N
    assert within_newton_tol(xm[:, both], xo[:, both])
    # the north star's bar for results: the FP64 lanes' node voltages to
    # 1e-9 relative (measured: 2e-13 V at most, every lane in the same ticks)
    same = both & (im == io)
    assert same.sum() >= 0.95 * N
    rel = np.abs(xm - xo)[:, same] / np.maximum(np.abs(xo[:, same]), 1e-3)
    assert rel.max() <= 1e-9


def test_mixed_opamp_transient_meas(engine):
    """The op-amp's whole lane (warm operating point, .tran 10n 1u, both
    .meas passes) with every Newton tick's linear step in FP32 plus FP64
    refinement, against the same lanes in FP64 (bitwise the oracle's,
    tests/test_gpu_va.py): the .meas values per lane, the accepted steps,
    the FP32 solves and the fallbacks."""
    N = 64
    say(f"case: mixed transient, psp103 op-amp, {N} lanes, .tran 10n 1u")
    inputs = W.opamp_inputs(N, seed=21)
    t = E.Template()
    eidx = lambda nm: E.param_index(E.PSP103, nm)
    inst = W.miller_opamp(t.add, eidx)
    t.freeze()
    g1 = E.Lanes(t, 1)
    g1.prepare()
    xn, sn, _, _, _ = g1.op_ladder()
    g = E.Lanes(t, N)
    for kind, name, slot, vals in W.opamp_lane_params(inputs, eidx):
        g.set_param(kind, inst[name], slot, vals)
    g.prepare()
    out = t.node("out")
    passes = [E.meas_extremum(out, 1), E.meas_cross(out, 0.95, rise=1)]
    x0 = np.repeat(xn, N, axis=1)
    g.upload_x(x0)
    r64, m64, _, _, ms64 = g.tran(E.tran_opts(1e-6, 10e-9, op_warm=True), passes)
    g.upload_x(x0)
    r32, m32, _, _, ms32 = g.tran(E.tran_opts(1e-6, 10e-9, op_warm=True, mixed_refine=10), passes)
    ok = (r64["status"] == 0) & (r32["status"] == 0)
    v64, v32 = m64[ok, :, 0], m32[ok, :, 0]
    # a pass that found nothing is NaN in both
    assert np.array_equal(np.isnan(v64), np.isnan(v32))
    fin = np.isfinite(v64)
    rel = np.abs(v32[fin] - v64[fin]) / np.abs(v64[fin])
    same_steps = int((r32["steps"] == r64["steps"]).sum())
    say(f"  completed {int(ok.sum())}/{N}, same accepted steps {same_steps}, FP32 solves per tick "
        f"{r32['fp32_solves'].sum() / max(r32['nr_iters'].sum(), 1):.2f}, fallbacks {int(r32['fp64_fallbacks'].sum())}, "
        f".meas max relative difference {rel.max():.2e} (median {np.median(rel):.2e}), "
        f"kernels {ms64:.1f} / {ms32:.1f} ms")
    assert ok.sum() == N
    assert (r64["fp32_solves"] == 0).all() and (r32["fp32_solves"] > 0).all()
    # the north star's bar: .meas values to 1e-9 relative (measured 3.2e-13,
    # every lane in the FP64 lanes' accepted steps)
    assert rel.max() <= 1e-9
