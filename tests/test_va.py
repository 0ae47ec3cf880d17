"""The Verilog-A compiler and the PSP103 model it compiles (CPU).

The generated C runs here through the oracle (oracle/va.c). What pins it:
central finite differences of every G and C entry (the derivative lanes),
an independent Python evaluation of the compiler's IR (the setup/evaluation
split, the relevance pruning and the emitted C, values only), KCL and
charge conservation over the device's rows, PSP's documented source/drain
symmetry. Value parity with the
reference's VerA-generated PSP103 is parity unpinned: VerA is not in the
reference tree and the reference cannot run here."""
import os

import numpy as np
import pytest

import oracle_lib as O
import va_interp as VI

VA = "/root/reference/models/psp103.va"
HAVE_SRC = os.path.exists(VA)
D, G, S, B, NOI, GP, SI, DI, BP, BI, BS, BD = range(12)


@pytest.fixture(scope="module")
def lib():
    return O.load()


def setup(lib, **named):
    n = lib.or_va_nparam(O.PSP103)
    u = np.full(n, np.nan)
    for k, v in named.items():
        i = lib.or_va_param_index(O.PSP103, k.encode())
        assert i >= 0, k
        u[i] = v
    p = np.zeros(lib.or_va_nprep(O.PSP103))
    rc = lib.or_va_setup(O.PSP103, O.f64(u)[1], O.f64(p)[1])
    return rc, p


def evaluate(lib, p, x):
    nu = lib.or_va_nu(O.PSP103)
    nj = sum(bin(lib.or_va_gmask(O.PSP103, r)).count("1") for r in range(nu))
    nc = sum(bin(lib.or_va_cmask(O.PSP103, r)).count("1") for r in range(nu))
    res, jac = np.zeros(nu), np.zeros(max(nj, 1))
    q, cj, site = np.zeros(nu), np.zeros(max(nc, 1)), np.zeros(max(1, lib.or_va_nsite(O.PSP103)))
    xa = np.ascontiguousarray(x, dtype=np.float64)
    lib.or_va_eval_tran(O.PSP103, O.f64(p)[1], O.f64(xa)[1], O.f64(res)[1], O.f64(jac)[1], O.f64(q)[1],
                        O.f64(cj)[1], O.f64(site)[1])
    return res, jac, q, cj, site


def dense(lib, flat, which):
    nu = lib.or_va_nu(O.PSP103)
    m = np.zeros((nu, nu))
    k = 0
    mask = lib.or_va_gmask if which == "g" else lib.or_va_cmask
    for r in range(nu):
        for c in range(nu):
            if mask(O.PSP103, r) >> c & 1:
                m[r, c] = flat[k]
                k += 1
    return m


def biases(rng, count, nu=12):
    x = rng.uniform(-0.3, 1.9, (count, nu))
    x[:, NOI] = 0.0
    return x


def test_generated_model_shape(lib):
    assert lib.or_va_nu(O.PSP103) == 12
    assert lib.or_va_nport(O.PSP103) == 4
    assert lib.or_va_nparam(O.PSP103) > 800
    for name in ("TYPE", "W", "L", "TOXO", "NSUBO", "RSH", "NRS", "DTEMP", "TRISE"):
        assert lib.or_va_param_index(O.PSP103, name.encode()) >= 0
    # DTEMP is an aliasparam of TRISE
    assert lib.or_va_param_index(O.PSP103, b"DTEMP") == lib.or_va_param_index(O.PSP103, b"TRISE")
    rc, _ = setup(lib, L=1e-12)  # below the declared range [1e-9, inf)
    assert rc == lib.or_va_param_index(O.PSP103, b"L") + 1


@pytest.mark.parametrize("ty", [1.0, -1.0])
def test_jacobian_matches_finite_differences(lib, ty):
    rc, p = setup(lib, TYPE=ty, W=2e-6, L=0.5e-6, RSH=10.0, NRS=1.0, NRD=1.0)
    assert rc == 0
    rng = np.random.default_rng(3)
    worst_g = worst_c = 0.0
    for x in biases(rng, 40) * ty:
        res, jac, q, cj, _ = evaluate(lib, p, x)
        Gm, Cm = dense(lib, jac, "g"), dense(lib, cj, "c")
        for c in range(12):
            h = 1e-6
            xp, xm = x.copy(), x.copy()
            xp[c] += h
            xm[c] -= h
            rp, _, qp, _, _ = evaluate(lib, p, xp)
            rm, _, qm, _, _ = evaluate(lib, p, xm)
            fd_g = (rp - rm) / (2 * h)
            fd_c = (qp - qm) / (2 * h)
            sg = np.abs(Gm).max() + 1e-15
            sc = np.abs(Cm).max() + 1e-30
            worst_g = max(worst_g, np.abs(fd_g - Gm[:, c]).max() / sg)
            worst_c = max(worst_c, np.abs(fd_c - Cm[:, c]).max() / sc)
            # entries outside the static pattern are identically zero
            gm = np.array([lib.or_va_gmask(O.PSP103, r) >> c & 1 for r in range(12)], bool)
            assert np.all(fd_g[~gm] == 0.0)
    # central differences of smooth bodies at h = 1e-6 V: O(h^2) truncation
    assert worst_g < 2e-6, worst_g
    assert worst_c < 2e-6, worst_c


def test_kcl_and_charge_conservation(lib):
    """With the noise node at 0 every current enters and leaves through the
    device's own rows, and every charge has its image."""
    rc, p = setup(lib, TYPE=1.0, W=2e-6, L=0.5e-6, RSH=10.0, NRS=1.0, NRD=1.0)
    rng = np.random.default_rng(4)
    for x in biases(rng, 60):
        res, _, q, _, _ = evaluate(lib, p, x)
        assert abs(res.sum()) <= 1e-12 * np.abs(res).sum() + 1e-25
        assert abs(q.sum()) <= 1e-12 * np.abs(q).sum() + 1e-35


def test_source_drain_symmetry(lib):
    """PSP is symmetric: exchanging the drain and source potentials
    exchanges their currents (the model's own sigVds swap)."""
    rc, p = setup(lib, TYPE=1.0, W=2e-6, L=0.5e-6)
    rng = np.random.default_rng(5)
    for _ in range(30):
        vd, vg, vs, vb = rng.uniform(0, 1.8, 4)
        vb = min(vb, vd, vs) - 0.1
        x = np.zeros(12)
        x[[D, DI]], x[[G, GP]], x[[S, SI]], x[[B, BP, BI, BS, BD]] = vd, vg, vs, vb
        y = x.copy()
        y[[D, DI]], y[[S, SI]] = vs, vd
        r1, _, _, _, _ = evaluate(lib, p, x)
        r2, _, _, _, _ = evaluate(lib, p, y)
        ids1 = r1[DI] + r1[D]
        ids2 = r2[SI] + r2[S]
        assert ids2 == pytest.approx(ids1, rel=1e-12, abs=1e-20)


def test_charge_only_evaluation_is_bitwise_the_transient_one(lib):
    """eval_charge (charge rows and sites, no derivative lanes: the lane
    engine's re-read of the charges at a solution) gives the same bits as
    eval_tran's charge rows and sites, both transistor types, biases across
    every region including reversed drain and source."""
    nu = lib.or_va_nu(O.PSP103)
    rng = np.random.default_rng(11)
    for ty, w, l in ((1, 2e-6, 0.5e-6), (-1, 8e-6, 1e-6), (1, 20e-6, 0.2e-6)):
        rc, p = setup(lib, TYPE=ty, W=w, L=l, RSH=10.0, NRS=1.0, NRD=1.0)
        assert rc == 0
        for x in biases(rng, 40) * ty:
            _, _, q, _, site = evaluate(lib, p, x)
            q2, s2 = np.zeros(nu), np.zeros(max(1, lib.or_va_nsite(O.PSP103)))
            xa = np.ascontiguousarray(x)
            lib.or_va_eval_charge(O.PSP103, O.f64(p)[1], O.f64(xa)[1], O.f64(q2)[1], O.f64(s2)[1])
            assert q.tobytes() == q2.tobytes()
            assert site.tobytes() == s2.tobytes()


@pytest.mark.skipif(not HAVE_SRC, reason="the Verilog-A source is not in this tree")
def test_emitted_code_matches_direct_evaluation(lib):
    """The generated setup + evaluation against a statement-by-statement
    walk of the same module in Python floats (libm transcendentals there,
    the oracle's own here: a few ulps)."""
    model = VI.load(VA)
    for user in (dict(TYPE=1, W=2e-6, L=0.5e-6, RSH=10.0, NRS=1.0, NRD=1.0),
                 dict(TYPE=-1, W=8e-6, L=1e-6, TRISE=60.0)):
        rc, p = setup(lib, **user)
        assert rc == 0
        it = VI.Interp(model, user)
        rng = np.random.default_rng(7)
        for x in biases(rng, 6) * user["TYPE"]:
            res, _, q, _, _ = evaluate(lib, p, x)
            r2, q2 = it.run(list(x))
            r2, q2 = np.array(r2), np.array(q2)
            np.testing.assert_allclose(res, r2, rtol=1e-9, atol=1e-9 * np.abs(r2).max() + 1e-24)
            np.testing.assert_allclose(q, q2, rtol=1e-9, atol=1e-9 * np.abs(q2).max() + 1e-35)


def test_opamp_operating_point_in_the_oracle(lib):
    """The two-stage op-amp as a unity-gain buffer: the oracle's CPU path
    reaches its operating point by gmin stepping; the output follows the
    input, KCL holds at every node and the supply delivers the bias
    branches' current."""
    from egspice_amd import workloads as W
    c = O.Circuit(lib)
    inst = W.miller_opamp(c.add, lambda nm: lib.or_va_param_index(O.PSP103, nm.encode()))
    c.freeze()
    L = O.Lanes(c, 1)
    L.prepare()
    x = np.zeros(c.n)
    for gmin in [1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10, 1e-11, 1e-12, 0.0]:
        xs, st, _ = L.newton_op(mode=1, x0=x, init_fix=False, gmin=gmin)
        assert st[0] & 1, gmin
        x = xs[:, 0]
    rows = {nm: c.node(nm) for nm in ("vdd", "inp", "out", "nbias", "tail", "n1", "n2")}
    assert abs(x[rows["out"]] - x[rows["inp"]]) < 5e-3
    assert 0.3 < x[rows["nbias"]] < 0.8
    rhs, _ = L.eval_lane(0, x)
    assert np.abs(rhs[1:]).max() < 1e-9
