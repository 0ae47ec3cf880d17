"""The operating point's ladder on the GPU against the oracle, bit for bit
(src/analysis/dc/op.zig:140-415): the homotopy Newton each stepping solve
runs (gmin at the ngspice stamps, the sources' factor, cold or warm, the
compiled kinds' $simparam("gmin")), and the whole ladder over lanes that
leave the plain rung at different rungs. Each case prints its name before
touching the GPU."""
import numpy as np
import pytest

import egspice_amd as E
import oracle_lib as O
from egspice_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def same_bits(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    return np.array_equal(a.view(np.uint64)[~both_nan], b.view(np.uint64)[~both_nan])


def chain_both(oracle, vin):
    t = E.Template()
    W.inverter_chain(t.add)
    t.freeze()
    g = E.Lanes(t, len(vin))
    g.set_param(E.VSOURCE, 1, E.VSOURCE_DC, vin)
    g.prepare()
    c = O.Circuit(oracle)
    W.inverter_chain(c.add)
    c.freeze()
    o = O.Lanes(c, len(vin))
    o.set_param(O.VSRC, 1, O.VSRC_DC, vin)
    o.prepare()
    return t, g, c, o


def opamp_both(oracle, n, seed):
    inputs = W.opamp_inputs(n, seed=seed)
    t = E.Template()
    eidx = lambda nm: E.param_index(E.PSP103, nm)
    inst = W.miller_opamp(t.add, eidx)
    t.freeze()
    g = E.Lanes(t, n)
    for kind, name, slot, vals in W.opamp_lane_params(inputs, eidx):
        g.set_param(kind, inst[name], slot, vals)
    g.prepare()
    oidx = lambda nm: oracle.or_va_param_index(O.PSP103, nm.encode())
    c = O.Circuit(oracle)
    oinst = W.miller_opamp(c.add, oidx)
    c.freeze()
    o = O.Lanes(c, n)
    for kind, name, slot, vals in W.opamp_lane_params(inputs, oidx):
        o.set_param(kind, oinst[name], slot, vals)
    o.prepare()
    return t, g, c, o


def test_homotopy_launch_bitwise(oracle, engine):
    """One homotopy launch over 24 lanes with mixed controls: shunts from 0
    to 1e-2 at the stamps, source factors in [0, 1], cold and warm starts,
    and inactive lanes, which keep their x and report {0, 0}."""
    N = 24
    say(f"case: homotopy Newton, inverter chain, {N} lanes, mixed controls")
    vin = np.linspace(0.0, 1.8, N)
    t, g, c, o = chain_both(oracle, vin)
    rng = np.random.default_rng(5)
    gshunt = np.where(rng.random(N) < 0.3, 0.0, 10.0 ** rng.uniform(-12, -2, N))
    srcfact = np.round(rng.random(N), 3)
    cold = rng.random(N) < 0.5
    active = rng.random(N) < 0.8
    flags = np.where(active, E.HOM_ACTIVE, 0) | np.where(cold, E.HOM_COLD, 0)
    # warm lanes start from the full-source answer scaled down
    xr, _, _, _ = g.newton_op()
    x0 = np.where(cold[None, :], 0.0, 0.5 * xr)
    g.upload_x(x0)
    xg, sg, ig, ms = g.newton_homotopy(gshunt, 1e-12, srcfact, flags, max_iter=100)
    say(f"  active {int(active.sum())}, converged {int((sg & 1).sum())}, kernel {ms:.3f} ms")
    for l in range(N):
        if not active[l]:
            assert sg[l] == 0 and ig[l] == 0
            assert same_bits(xg[:, l], x0[:, l])
            continue
        xo, so, io = o.newton_hom(l, x0[:, l], gshunt[l], srcfact[l], init_fix=bool(cold[l]), max_iter=100)
        assert sg[l] == so and ig[l] == io, (l, sg[l], so, ig[l], io)
        assert same_bits(xg[:, l], xo), l


@pytest.mark.parametrize("plain_iter,ladder_iter,n,rungs", [
    (2, 100, 8, None),         # every lane by gmin stepping
    (6, 6, 4, [0, 3, 2, 2]),   # none, source stepping, gmin stepping
    (8, 3, 4, [0, 3, 3, 0]),
])
def test_ladder_chain_bitwise(oracle, engine, plain_iter, ladder_iter, n, rungs):
    say(f"case: ladder, inverter chain, {n} lanes, plain {plain_iter} / stepping {ladder_iter} iterations")
    vin = np.linspace(0.0, 1.8, n)
    t, g, c, o = chain_both(oracle, vin)
    xg, sg, ig, rg, ms = g.op_ladder(max_iter=plain_iter, ladder_max_iter=ladder_iter)
    xo, so, io, ro = o.op_ladder(max_iter=plain_iter, ladder_max_iter=ladder_iter)
    say(f"  rungs {list(rg)} iterations {list(ig)} kernels {ms:.3f} ms")
    if rungs is not None:
        assert list(ro) == rungs
    else:
        assert (ro == 2).all()
    assert np.array_equal(rg, ro) and np.array_equal(sg, so) and np.array_equal(ig, io)
    assert same_bits(xg, xo)


def test_ladder_opamp_bitwise(oracle, engine):
    """The psp103 op-amp's cold operating point fails the plain rung and
    takes the gmin descent, whose solves publish their gmin to PSP103's
    $simparam("gmin") (the engine writes it into the prepared constants,
    the oracle reruns the setup)."""
    N = 8
    say(f"case: ladder, psp103 op-amp Monte Carlo, {N} lanes from a cold start")
    t, g, c, o = opamp_both(oracle, N, seed=3)
    xg, sg, ig, rg, ms = g.op_ladder()
    xo, so, io, ro = o.op_ladder()
    say(f"  rungs {list(rg)} iterations {list(ig)} kernels {ms:.3f} ms")
    assert (so == 1).all() and (ro == 2).all()
    assert np.array_equal(rg, ro) and np.array_equal(sg, so) and np.array_equal(ig, io)
    assert same_bits(xg, xo)
    # the lanes' model gmin is the deck's again: a plain warm Newton from
    # the answer is bitwise the oracle's
    g.upload_x(xg)
    xg2, sg2, ig2, _ = g.newton(init_fix=False)
    xo2, so2, io2 = o.newton_op(mode=1, x0=xo.reshape(-1), init_fix=False)
    assert np.array_equal(sg2, so2) and np.array_equal(ig2, io2) and same_bits(xg2, xo2)


def sweep_values(start, step, npts):
    v, out = start, []
    for k in range(npts):
        if k:
            v += step
        out.append(v)
    return np.array(out)


@pytest.mark.parametrize("warm_iter", [100, 2])
def test_dc_sweep_chain_bitwise(oracle, engine, warm_iter):
    """The `.dc` sweep over 64 Monte Carlo lanes of the inverter chain (VTO
    spread), 0-1.8 V in 19 points: warm from the limited predictor, and
    with a 2-iteration warm budget through the cold ladder at most points;
    x and convergence bitwise the oracle's."""
    N = 64
    say(f"case: .dc sweep, inverter chain, {N} lanes, 19 points, warm budget {warm_iter}")
    vin = sweep_values(0.0, 0.1, 19)
    vto = np.linspace(0.6, 0.8, N)
    t = E.Template()
    W.inverter_chain(t.add)
    t.freeze()
    g = E.Lanes(t, N)
    c = O.Circuit(oracle)
    W.inverter_chain(c.add)
    c.freeze()
    o = O.Lanes(c, N)
    for i in range(0, 10, 2):
        g.set_param(E.MOS1, i, E.MOS1_PARAMS["VTO"], vto)
        o.set_param(O.MOS1, i, O.M1["VTO"], vto)
    g.prepare()
    o.prepare()
    xg, cg, ms = g.dc_sweep(E.VSOURCE, 1, vin, ladder_max_iter=warm_iter)
    xo, co = o.dc_sweep_lanes(O.VSRC, 1, vin, ladder_max_iter=warm_iter)
    say(f"  converged {int(cg.sum())}/{cg.size}, kernels {ms:.3f} ms")
    assert co.all()
    assert np.array_equal(cg, co)
    assert same_bits(xg, xo)


def test_dc_sweep_diode_golden(oracle, engine):
    """docs/examples/diode_iv.sp on 4 lanes: the reference's printed i(v1)
    at 10 and 20 mV (docs/learn/analyses.md:88-93) and bitwise the oracle."""
    say("case: .dc sweep, diode_iv.sp, 4 lanes, 81 points")
    vin = sweep_values(0.0, 0.01, 81)
    t = E.Template()
    t.add(E.VSOURCE, ["a", "0"], {E.VSOURCE_DC: 0.0})
    t.add(E.DIODE, ["a", "0"], {E.DIODE_PARAMS["IS"]: 1e-14, E.DIODE_PARAMS["N"]: 1.0})
    t.freeze()
    g = E.Lanes(t, 4)
    g.prepare()
    c = O.Circuit(oracle)
    c.add(O.VSRC, ["a", "0"], {O.VSRC_DC: 0.0})
    c.add(O.DIODE, ["a", "0"], {O.DIO["IS"]: 1e-14, O.DIO["N"]: 1.0})
    c.freeze()
    o = O.Lanes(c, 4)
    o.prepare()
    xg, cg, _ = g.dc_sweep(E.VSOURCE, 0, vin)
    xo, co = o.dc_sweep_lanes(O.VSRC, 0, vin)
    assert cg.all() and np.array_equal(cg, co) and same_bits(xg, xo)
    for k, v in {1: -1.386624089968383e-14, 2: -3.166792969038269e-14}.items():
        assert (np.abs(xg[k, 2] / v - 1) <= 4e-16).all()
