"""BSIM4 (models/bsim4va.va) compiled by the engine's Verilog-A compiler (CPU).

The generated C runs here through the oracle (oracle/va.c). What pins it:
central finite differences of every G and C entry, KCL and charge
conservation over the device's rows, an independent Python evaluation of
the compiler's IR, the charge-only evaluation against the transient one,
and the compiled limiter (the model's $limit group) against the hand-written
mos1 limiter, which restates the same mos1load.c rung, on prepared constants
chosen to give both the same threshold, thermal voltage and critical
voltage. Value parity with the reference's VerA-generated BSIM4 is parity
unpinned: VerA is not in the reference tree and the reference cannot run
here."""
import os

import numpy as np
import pytest

import oracle_lib as O
import va_interp as VI

VA = "/root/reference/models/bsim4va.va"
HAVE_SRC = os.path.exists(VA)
K = O.BSIM4
D, G, S, B, DI, SI, GI, GM, BI, SB, DB = range(11)
# every internal unknown on its port (rdsmod = rgatemod = rbodymod = 0)
PORT_OF = {DI: D, SI: S, GI: G, GM: G, BI: B, SB: B, DB: B}


@pytest.fixture(scope="module")
def lib():
    return O.load()


def setup(lib, **named):
    n = lib.or_va_nparam(K)
    u = np.full(n, np.nan)
    for k, v in named.items():
        i = lib.or_va_param_index(K, k.encode())
        assert i >= 0, k
        u[i] = v
    p = np.zeros(lib.or_va_nprep(K))
    rc = lib.or_va_setup(K, O.f64(u)[1], O.f64(p)[1])
    return rc, p


def evaluate(lib, p, x):
    nu = lib.or_va_nu(K)
    nj = sum(bin(lib.or_va_gmask(K, r)).count("1") for r in range(nu))
    nc = sum(bin(lib.or_va_cmask(K, r)).count("1") for r in range(nu))
    res, jac = np.zeros(nu), np.zeros(max(nj, 1))
    q, cj, site = np.zeros(nu), np.zeros(max(nc, 1)), np.zeros(max(1, lib.or_va_nsite(K)))
    xa = np.ascontiguousarray(x, dtype=np.float64)
    lib.or_va_eval_tran(K, O.f64(p)[1], O.f64(xa)[1], O.f64(res)[1], O.f64(jac)[1], O.f64(q)[1], O.f64(cj)[1],
                        O.f64(site)[1])
    return res, jac, q, cj, site


def dense(lib, flat, which):
    nu = lib.or_va_nu(K)
    m = np.zeros((nu, nu))
    k = 0
    mask = lib.or_va_gmask if which == "g" else lib.or_va_cmask
    for r in range(nu):
        for c in range(nu):
            if mask(K, r) >> c & 1:
                m[r, c] = flat[k]
                k += 1
    return m


def biases(rng, count, ty):
    """Port potentials of an inverter's transistor (source at 0 or 1.8 V
    for the type), internal unknowns on their ports plus a few mV."""
    out = []
    for _ in range(count):
        x = np.zeros(11)
        x[S] = 0.0 if ty > 0 else 1.8
        x[B] = x[S] + ty * rng.uniform(-0.3, 0.0)
        x[G] = rng.uniform(0, 1.8)
        x[D] = rng.uniform(0, 1.8)
        for u, port in PORT_OF.items():
            x[u] = x[port] + rng.uniform(-3e-3, 3e-3)
        out.append(x)
    return out


NMOS = dict(TYPE=1.0, W=1e-6, L=0.1e-6)
PMOS = dict(TYPE=-1.0, W=2e-6, L=0.1e-6)


def test_bsim4_model_shape(lib):
    assert lib.or_va_nu(K) == 11
    assert lib.or_va_nport(K) == 4
    assert lib.or_va_nparam(K) > 800
    for name in ("TYPE", "W", "L", "VTH0", "TOXE", "NF", "RGATEMOD", "RBODYMOD", "RDSMOD"):
        assert lib.or_va_param_index(K, name.encode()) >= 0, name
    # the limiter moves the internal gate, drain and bulk, and every unknown
    # a collapse can join to them
    assert lib.or_va_limw(K) == sum(1 << u for u in (D, G, B, DI, GI, GM, BI, SB, DB))
    rc, _ = setup(lib, L=-1.0)
    assert rc == lib.or_va_param_index(K, b"L") + 1


@pytest.mark.parametrize("dev", [NMOS, PMOS], ids=["nmos", "pmos"])
def test_bsim4_jacobian_matches_finite_differences(lib, dev):
    rc, p = setup(lib, **dev)
    assert rc == 0
    rng = np.random.default_rng(3)
    worst_g = worst_c = 0.0
    for x in biases(rng, 25, dev["TYPE"]):
        _, jac, _, cj, _ = evaluate(lib, p, x)
        Gm, Cm = dense(lib, jac, "g"), dense(lib, cj, "c")
        for c in range(11):
            h = 1e-6
            xp, xm = x.copy(), x.copy()
            xp[c] += h
            xm[c] -= h
            rp, _, qp, _, _ = evaluate(lib, p, xp)
            rm, _, qm, _, _ = evaluate(lib, p, xm)
            fd_g = (rp - rm) / (2 * h)
            fd_c = (qp - qm) / (2 * h)
            worst_g = max(worst_g, np.abs(fd_g - Gm[:, c]).max() / (np.abs(Gm).max() + 1e-15))
            worst_c = max(worst_c, np.abs(fd_c - Cm[:, c]).max() / (np.abs(Cm).max() + 1e-30))
            gm = np.array([lib.or_va_gmask(K, r) >> c & 1 for r in range(11)], bool)
            assert np.all(fd_g[~gm] == 0.0)
    assert worst_g < 1e-5, worst_g
    assert worst_c < 1e-5, worst_c


def test_bsim4_kcl_and_charge_conservation(lib):
    for dev in (NMOS, PMOS):
        rc, p = setup(lib, **dev)
        rng = np.random.default_rng(4)
        for x in biases(rng, 40, dev["TYPE"]):
            res, _, q, _, _ = evaluate(lib, p, x)
            assert abs(res.sum()) <= 1e-12 * np.abs(res).sum() + 1e-25
            assert abs(q.sum()) <= 1e-12 * np.abs(q).sum() + 1e-35


def test_bsim4_charge_only_evaluation_is_bitwise_the_transient_one(lib):
    nu = lib.or_va_nu(K)
    for dev in (NMOS, PMOS):
        rc, p = setup(lib, **dev)
        rng = np.random.default_rng(11)
        for x in biases(rng, 40, dev["TYPE"]):
            _, _, q, _, site = evaluate(lib, p, x)
            q2, s2 = np.zeros(nu), np.zeros(max(1, lib.or_va_nsite(K)))
            lib.or_va_eval_charge(K, O.f64(p)[1], O.f64(np.ascontiguousarray(x))[1], O.f64(q2)[1], O.f64(s2)[1])
            assert q.tobytes() == q2.tobytes()
            assert site.tobytes() == s2.tobytes()


@pytest.mark.skipif(not HAVE_SRC, reason="the Verilog-A source is not in this tree")
def test_bsim4_emitted_code_matches_direct_evaluation(lib):
    """The generated setup + evaluation against a statement-by-statement
    walk of the module in Python floats (libm there, the oracle's own
    elementary functions here: a few ulps, amplified by cancellation in
    the subthreshold and junction currents)."""
    model = VI.load(VA)
    for dev in (NMOS, PMOS):
        rc, p = setup(lib, **dev)
        assert rc == 0
        it = VI.Interp(model, dev)
        rng = np.random.default_rng(7)
        for x in biases(rng, 4, dev["TYPE"]):
            res, _, q, _, _ = evaluate(lib, p, x)
            r2, q2 = it.run(list(x))
            r2, q2 = np.array(r2), np.array(q2)
            np.testing.assert_allclose(res, r2, rtol=1e-8, atol=1e-8 * np.abs(r2).max() + 1e-24)
            np.testing.assert_allclose(q, q2, rtol=1e-8, atol=1e-8 * np.abs(q2).max() + 1e-35)


M1P = dict(TYPE=0, VT=1, VBIT=2, PHIT=3, GAMMA=4, VCRITS=10, VCRITD=11)


@pytest.mark.parametrize("dev", [NMOS, PMOS], ids=["nmos", "pmos"])
def test_bsim4_limiter_matches_the_mos1_rung(lib, dev):
    """bsim4va's $limit group (fetlimds on both gate legs, limvds on the
    channel, pnjlimds on both bulk legs) against the hand-written mos1
    limiter, which restates mos1load.c's rung: mos1 constants with GAMMA = 0
    and VBIT * type = the group's threshold, VT and VCRIT its thermal and
    critical voltages. The limited gate, drain and bulk images and pnjlim's
    check agree bit for bit over jumps in every region; the seeds are the
    MODEINITJCT values the model names (vgs = type*vth0 + 0.1, vds = 0.1,
    vbs = 0)."""
    rc, p = setup(lib, **dev)
    assert rc == 0
    ty = dev["TYPE"]
    la = np.zeros(16)
    lib.or_va_limit_args(K, O.f64(p)[1], O.f64(np.zeros(11))[1], O.f64(la)[1])
    # fetlimds (von, type, seed), fetlimds (von, type), limvds (type, seed),
    # pnjlimds (vt, vcrit, type, seed), pnjlimds (vt, vcrit, type)
    assert la[1] == ty and la[4] == ty and la[5] == ty and la[9] == ty and la[13] == ty
    assert la[0] == la[3] and la[7] == la[11] and la[8] == la[12]
    assert 0.05 < abs(la[0]) < 1.0 and la[2] == la[0] + 0.1 and la[6] == 0.1 and la[10] == 0.0
    # the thermal voltage at 300.15 K and the junction's critical voltage
    kt = 1.38064852e-23 * 300.15 / 1.6021766208e-19
    assert la[7] == pytest.approx(kt, rel=1e-15)
    assert la[8] == pytest.approx(kt * np.log(kt / (np.sqrt(2.0) * 1e-14)), rel=1e-13)
    m = np.zeros(40)
    m[M1P["TYPE"]] = ty
    m[M1P["VT"]] = la[7]
    m[M1P["VBIT"]] = la[0] * ty
    m[M1P["PHIT"]] = 0.6
    m[M1P["VCRITS"]] = la[8]
    m[M1P["VCRITD"]] = la[12]
    rng = np.random.default_rng(13)
    checks = 0
    for trial in range(2000):
        c = rng.uniform(-1, 3, 11) * ty
        o = rng.uniform(-1, 3, 11) * ty
        if trial % 4 == 0:
            o = c + rng.normal(0, 0.05, 11)  # small steps: the limiter leaves them
        lc = np.zeros(11)
        chk_c = lib.or_va_limit(K, O.f64(p)[1], O.f64(c)[1], O.f64(o)[1], O.f64(lc)[1])
        # mos1's unknowns (d, g, s, b) are the limiter's (di, gi, si, bi)
        cm = np.ascontiguousarray([c[DI], c[GI], c[SI], c[BI]])
        om = np.ascontiguousarray([o[DI], o[GI], o[SI], o[BI]])
        lm = np.zeros(4)
        chk_m = lib.or_limit_mos1(O.f64(m)[1], O.f64(cm)[1], O.f64(om)[1], O.f64(lm)[1])
        assert chk_c == chk_m
        checks += chk_c
        assert np.all(np.isfinite(lc)) and np.all(np.isfinite(lm))
        assert np.array([lc[DI], lc[GI], lc[SI], lc[BI]]).tobytes() == lm.tobytes()
        # collapsed onto the limited gate, drain and bulk: their images
        assert lc[G] == lc[GM] == lc[GI] and lc[D] == lc[DI] and lc[B] == lc[SB] == lc[DB] == lc[BI]
        # the source side is left as it is
        assert lc[S] == c[S] and lc[SI] == c[SI]
    assert 100 < checks < 1900
    # seeds
    cur = rng.uniform(0, 1.8, 11)
    lim = np.zeros(11)
    lib.or_va_seed(K, O.f64(p)[1], O.f64(cur)[1], O.f64(lim)[1])
    s0 = cur[SI]
    assert lim[DI] == s0 + ty * 0.1
    assert lim[GI] == s0 + ty * la[2]
    assert lim[BI] == s0 + ty * 0.0
