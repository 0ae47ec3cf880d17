"""GPU parity of the lane engine against the CPU oracle, through the C ABI.

Integer and index work must match bit for bit; the floating-point lane
results must too, since the engine runs the oracle's operation order with
contraction off (tolerance: 0 ulp; a NaN matches any NaN). Each case prints its name and sizes
before touching the GPU.
"""
import sys

import numpy as np
import pytest

import egspice_amd as E
from egspice_amd import workloads as W
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build_both(oracle, builder, *args):
    t = E.Template()
    builder(t.add, *args)
    t.freeze()
    c = O.Circuit(oracle)
    builder(c.add, *args)
    c.freeze()
    assert (t.col_ptr == c.col_ptr).all() and (t.row_idx == c.row_idx).all()
    return t, c


def bits_equal(a, b):
    """Bit for bit, except that any NaN equals any NaN: IEEE 754 leaves the
    sign and payload of an operation's NaN to the platform (x86 returns the
    negative default NaN, gfx950 a positive one), and both only arise in
    lanes that went singular."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    same = a.view(np.uint64) == b.view(np.uint64)
    return bool(np.all(same | (np.isnan(a) & np.isnan(b))))


def test_config1_dc_sweep_lanes_bitwise(oracle, engine):
    say("case: inverter chain 5 stages, 64 lanes, cold OP, engine vs oracle CPU path")
    t, c = build_both(oracle, W.inverter_chain)
    vin = W.dc_sweep_points(64)
    N = len(vin)
    gl = E.Lanes(t, N)
    gl.set_param(E.VSOURCE, 1, E.VSOURCE_DC, vin)
    gl.prepare()
    xg, sg, ig, ms = gl.newton_op()
    ol = O.Lanes(c, N)
    ol.set_param(O.VSRC, 1, O.VSRC_DC, vin)
    ol.prepare()
    xo, so, io = ol.newton_op(mode=1)
    say(f"  converged gpu {int((sg & 1).sum())} oracle {int((so & 1).sum())} kernel {ms:.3f} ms")
    assert (so & 1).all()
    assert (sg == so).all()
    assert (ig == io).all()
    assert bits_equal(xg, xo)


@pytest.mark.parametrize("seed", [1, 2])
def test_mc_lanes_bitwise(oracle, engine, seed):
    N = 1000
    say(f"case: inverter chain Monte Carlo, {N} lanes, seed {seed}")
    t, c = build_both(oracle, W.inverter_chain)
    rng = np.random.default_rng(seed)
    vin = rng.uniform(0.0, 1.8, N)
    vto_n = 0.7 + 0.015 * rng.standard_normal(N)
    tox = 1e-7 * (1 + 0.02 * rng.standard_normal(N))
    gl = E.Lanes(t, N)
    ol = O.Lanes(c, N)
    for L, kv, km in ((gl, E.VSOURCE, E.MOS1), (ol, O.VSRC, O.MOS1)):
        L.set_param(kv, 1, 0, vin)
        L.set_param(km, 0, O.M1["VTO"], vto_n)
        L.set_param(km, 2, O.M1["TOX"], tox)
        L.prepare()
    xg, sg, ig, ms = gl.newton_op()
    xo, so, io = ol.newton_op(mode=1)
    say(f"  converged gpu {int((sg & 1).sum())} oracle {int((so & 1).sum())} kernel {ms:.3f} ms")
    assert (sg == so).all()
    assert (ig == io).all()
    assert bits_equal(xg, xo)


def random_system(rng, n, density):
    """sparse_lu.zig:1380-1410 TestSystem: full diagonal, off-diagonals at
    `density`, diagonal pushed past its column sum except a zero (branch
    row) diagonal 15% of the time."""
    col_ptr, row_idx, vals = [0], [], []
    for j in range(n):
        rows = [i for i in range(n) if i == j or rng.random() < density]
        col = [2 * rng.random() - 1 for _ in rows]
        d = rows.index(j)
        s = sum(abs(v) for i, v in enumerate(col) if i != d)
        col[d] = 0.0 if (s > 0 and rng.random() < 0.15) else col[d] + np.copysign(s, col[d])
        row_idx += rows
        vals += col
        col_ptr.append(len(row_idx))
    return np.array(col_ptr, np.uint32), np.array(row_idx, np.uint32), np.array(vals)


def test_lead_tape_replay_bitwise(oracle, engine):
    """lane.zig:811-908: every lane's replay of the lead tape is bitwise the
    oracle's body, failing exactly where it fails."""
    rng = np.random.default_rng(0x1A7E51)
    W_ = 64
    checked = fails = 0
    for trial in range(40):
        n = 2 + trial % 20
        cp, ri, v = random_system(rng, n, 0.3)
        try:
            tab, idx = E.tab_build(n, cp, ri, v)
        except E.EngineError:
            continue
        # oracle tape from the same factor
        s = oracle.or_solver_new(n, O.u32(cp)[1], O.u32(ri)[1])
        if oracle.or_solver_factor(s, O.f64(v)[1]) != 0:
            oracle.or_solver_free(s)
            continue
        import ctypes as C
        otab = O.LaneTab()
        ln = C.c_size_t()
        oidx_p = oracle.or_build_tab(oracle.or_solver_lu(s), O.u32(cp)[1], oracle.or_solver_q(s), 1e-12, 1e-3,
                                     C.byref(otab), C.byref(ln))
        oidx = np.ctypeslib.as_array(oidx_p, (ln.value,)).copy()
        oracle.free_(C.cast(oidx_p, C.c_void_p))
        oracle.or_solver_free(s)
        assert np.array_equal(oidx, idx), f"tape words differ at n={n}"
        nnz = len(v)
        a = np.repeat(v[:, None], W_, axis=1)
        for l in range(W_):
            kind = rng.integers(6)
            if kind == 0:
                a[rng.integers(nnz), l] = 4 * rng.random() - 2
            elif kind == 1:
                col = rng.integers(n)
                a[cp[col]:cp[col + 1], l] = 0
            else:
                a[:, l] = v * (0.95 + 0.1 * rng.random(nnz))
        rhs = 2 * rng.random((n, W_)) - 1
        say(f"case: lead tape n={n} nnz={nnz} lanes={W_}")
        dx, fail, _ = E.tab_refactor_solve(tab, idx, a, rhs)
        lx = np.zeros((max(otab.nnz_l, 1), W_))
        ux = np.zeros((max(otab.nnz_u, 1), W_))
        ud, w, y, dxo = (np.zeros((n, W_)) for _ in range(4))
        for l in range(W_):
            bad = oracle.or_lane_refactor(C.byref(otab), O.u32(oidx)[1], O.f64(a)[1], lx.ctypes.data_as(C.POINTER(C.c_double)),
                                          ux.ctypes.data_as(C.POINTER(C.c_double)), ud.ctypes.data_as(C.POINTER(C.c_double)),
                                          w.ctypes.data_as(C.POINTER(C.c_double)), W_, l)
            assert int(bad) == int(fail[l])
            if bad:
                fails += 1
                continue
            oracle.or_lane_solve_neg(C.byref(otab), O.u32(oidx)[1], lx.ctypes.data_as(C.POINTER(C.c_double)),
                                     ux.ctypes.data_as(C.POINTER(C.c_double)), ud.ctypes.data_as(C.POINTER(C.c_double)),
                                     O.f64(rhs)[1], y.ctypes.data_as(C.POINTER(C.c_double)),
                                     dxo.ctypes.data_as(C.POINTER(C.c_double)), W_, l)
            assert bits_equal(dx[:, l], dxo[:, l])
            checked += 1
    assert checked > 0 and fails > 0


def run_pair(oracle, builder, N, params, args=(), **kw):
    t, c = build_both(oracle, builder, *args)
    gl, ol = E.Lanes(t, N), O.Lanes(c, N)
    for L in (gl, ol):
        for kind, inst, k, vals in params:
            L.set_param(kind, inst, k, vals)
        L.prepare()
    xg, sg, ig, ms = gl.newton_op(**kw)
    xo, so, io = ol.newton_op(mode=1, **kw)
    say(f"  n={t.n} converged gpu {int((sg & 1).sum())} oracle {int((so & 1).sum())} kernel {ms:.3f} ms")
    assert (sg == so).all()
    assert (ig == io).all()
    assert bits_equal(xg, xo)
    return so


@pytest.mark.parametrize("stages", [1, 12, 30, 40, 59, 60])
def test_chain_sizes_bitwise(oracle, engine, stages):
    """n = stages + 5 unknowns. Up to 40 stages the lanes converge; at 59
    and 60 (n = 64 and 65) a stage with both devices cut off leaves a node
    without conductance (lambda, junction leakage and gmin all zero), and
    every lane must go singular exactly where the oracle's does."""
    N = 64
    say(f"case: inverter chain {stages} stages, {N} lanes")
    rng = np.random.default_rng(stages)
    so = run_pair(oracle, W.inverter_chain, N, [(E.VSOURCE, 1, E.VSOURCE_DC, rng.uniform(0, 1.8, N))], (stages,))
    if stages <= 40:
        assert (so & 1).mean() > 0.9


def test_tridiagonal_thomas_lanes_bitwise(oracle, engine):
    from test_host_bodies import diode_only
    say("case: diode deck (tridiagonal, Thomas engine), 81 lanes")
    so = run_pair(oracle, diode_only, 81, [(E.VSOURCE, 0, E.VSOURCE_DC, np.arange(81) * 0.01)])
    assert (so & 1).all()


def test_mixed_kinds_lanes_bitwise(oracle, engine):
    from test_host_bodies import mixed
    say("case: divider, diodes, current sink, capacitor; 256 lanes")
    rng = np.random.default_rng(9)
    N = 256
    so = run_pair(oracle, mixed, N, [(E.VSOURCE, 0, E.VSOURCE_DC, rng.uniform(-2, 8, N)),
                                     (E.ISOURCE, 0, 0, rng.uniform(-2e-3, 2e-3, N)),
                                     (E.DIODE, 1, O.DIO["IS"], 10 ** rng.uniform(-16, -12, N))])
    assert (so & 1).all()


@pytest.mark.parametrize("kw", [{"max_iter": 3}, {"max_iter": 7, "init_fix": False}, {"init_fix": False},
                                {"gmin": 1e-12}])
def test_newton_options_bitwise(oracle, engine, kw):
    say(f"case: inverter chain, 64 lanes, options {kw}")
    run_pair(oracle, W.inverter_chain, 64, [(E.VSOURCE, 1, E.VSOURCE_DC, W.dc_sweep_points(64))], **kw)


def test_monte_carlo_wide_bitwise(oracle, engine):
    """Many lanes per CU and spread parameters: lambda, body effect and
    junction capacitances on, the 4096-lane case."""
    N = 4096
    say(f"case: inverter chain Monte Carlo, {N} lanes, lambda and gamma spread")
    rng = np.random.default_rng(77)
    params = [(E.VSOURCE, 1, 0, rng.uniform(0, 1.8, N)),
              (E.MOS1, 0, O.M1["VTO"], 0.7 + 0.015 * rng.standard_normal(N)),
              (E.MOS1, 1, O.M1["VTO"], -0.7 + 0.015 * rng.standard_normal(N)),
              (E.MOS1, 4, O.M1["LAMBDA"], rng.uniform(0, 0.05, N)),
              (E.MOS1, 6, O.M1["GAMMA"], rng.uniform(0, 0.5, N)),
              (E.MOS1, 2, O.M1["TOX"], 1e-7 * (1 + 0.02 * rng.standard_normal(N)))]
    so = run_pair(oracle, W.inverter_chain, N, params)
    assert (so & 1).mean() > 0.95


def test_sampler_drawn_lanes_bitwise(oracle, engine):
    """Per-lane draws from the reference's sampler, as a Monte Carlo sweep
    feeds them: every MOSFET's VTO ~ agauss(nom, 45 mV, 3) and KP ~
    gauss(nom, 2 %, 1), keyed (seed, trial, site) with one site per device
    parameter; the input swept uniformly."""
    N = 2048
    say(f"case: inverter chain, {N} trials drawn by the Monte Carlo sampler (seed 7)")
    seed = 7
    params = [(E.VSOURCE, 1, 0, E.mc_draw(seed, E.site_key(1, 0), 1, N, E.MC_AUNIF, 0.9, 0.9))]
    for k in range(10):
        nom_vto = 0.7 if k % 2 == 0 else -0.7
        nom_kp = 1e-4 if k % 2 == 0 else 4e-5
        params.append((E.MOS1, k, O.M1["VTO"], E.mc_draw(seed, E.site_key(100 + k, 1), 1, N, E.MC_AGAUSS,
                                                          nom_vto, 0.045, 3)))
        params.append((E.MOS1, k, O.M1["KP"], E.mc_draw(seed, E.site_key(100 + k, 2), 1, N, E.MC_GAUSS,
                                                         nom_kp, 0.02, 1)))
    so = run_pair(oracle, W.inverter_chain, N, params)
    assert (so & 1).mean() > 0.95


def divider3(add):
    """V1 a 0; R1 a b; R2 b 0; R3 a 0: one node column holds the source's
    branch entry, so its pivot depends on the conductances' scale."""
    add(E.VSOURCE, ["a", "0"], {0: 1.0})
    add(E.RESISTOR, ["a", "b"], {0: 1.0})
    add(E.RESISTOR, ["b", "0"], {0: 1.0})
    add(E.RESISTOR, ["a", "0"], {0: 1.0})


def first_factor_pair(oracle, builder, N, params, args=()):
    t, c = build_both(oracle, builder, *args)
    gl, ol = E.Lanes(t, N), O.Lanes(c, N)
    for L in (gl, ol):
        for kind, inst, k, vals in params:
            L.set_param(kind, inst, k, vals)
        L.prepare()
    xg, sg, ig, _ = gl.newton_op()
    xo, so, io = ol.newton_op(mode=1)
    assert (sg == so).all() and (ig == io).all() and bits_equal(xg, xo)
    return gl.first_factors(), so


def test_predicted_first_factor_hits_on_the_sweep(oracle, engine):
    """Every lane of the 64-point inverter-chain sweep takes the template's
    predicted first factor, and the lanes stay bitwise the oracle's (whose
    full factor is the serial one)."""
    say("case: predicted first factor, inverter chain, 64 lanes")
    (pred, serial), so = first_factor_pair(oracle, W.inverter_chain, 64,
                                           [(E.VSOURCE, 1, E.VSOURCE_DC, W.dc_sweep_points(64))])
    say(f"  predicted {pred}, serial {serial}")
    assert (pred, serial) == (64, 0)
    assert (so & 1).all()


def test_predicted_first_factor_falls_back(oracle, engine):
    """Conductances over nine decades per lane, and on every fourth lane R1
    and R2 open (infinite), which leaves node b's column without a nonzero
    candidate: those lanes' first factors cannot follow the prediction
    (taken from lane 0, all finite) and run the serial factor, which decides
    the void column; the others take the prediction. Every lane stays
    bitwise the oracle's."""
    N = 64
    rng = np.random.default_rng(31)
    vals = [np.concatenate([[1.0], 10 ** rng.uniform(0, 9, N - 1)]) for _ in range(3)]
    open_b = np.arange(N) % 4 == 1
    vals[0][open_b] = np.inf
    vals[1][open_b] = np.inf
    say(f"case: predicted first factor with fallbacks, 3-resistor divider, {N} lanes, {int(open_b.sum())} with b open")
    (pred, serial), so = first_factor_pair(oracle, divider3, N,
                                           [(E.RESISTOR, k, 0, vals[k]) for k in range(3)])
    say(f"  predicted {pred}, serial {serial}, converged {int((so & 1).sum())}")
    assert serial == int(open_b.sum()) and pred == N - serial
    assert (so[~open_b] & 1).all()
