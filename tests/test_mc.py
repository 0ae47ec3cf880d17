"""The Monte Carlo sampler (per-lane parameter draws), engine and oracle.

Pinned without retyping its formulas:
  - site_key is splitmix64's finalizer: site_key(0, (k-1)·γ) must give
    splitmix64's published outputs k = 1..4 from state 0;
  - the normal quantile against Python's statistics.NormalDist within
    Acklam's stated 1.2e-9 relative error, and the reference's own tests
    (src/core/numerics.zig:520-545): Φ⁻¹(0.5) = 0, Φ⁻¹(0.975) =
    1.959963984540054 to 1e-8, odd symmetry, monotone, NaN at 0 and 1;
  - TrialDraw's tests (src/frontend/variants.zig:1917-1933): `limit` lands
    on nom ± abs exactly, `aunif` stays inside its band, a multiplier keeps
    the largest deviation, a nominal-only call returns nom;
  - statistics of many trials: agauss(nom, abs, sigmas) has mean nom and
    sigma abs/sigmas; unif is uniform on nom·(1 ± rvar);
and the engine's draws equal the oracle's bit for bit."""
import ctypes as C
import statistics

import numpy as np
import pytest

import egspice_amd as E
import oracle_lib as O

GAMMA = 0x9e3779b97f4a7c15


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))
SPLITMIX64_FROM_0 = [0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f, 0xf88bb8a8724c81ec]


def test_site_key_is_splitmix64(oracle, engine):
    for k, want in enumerate(SPLITMIX64_FROM_0):
        salt = (k * GAMMA) % 2**64
        assert E.site_key(0, salt) == want
        assert oracle.or_site_key(0, salt) == want


def test_normal_quantile(oracle):
    q = oracle.or_normal_quantile
    assert abs(q(0.5)) <= 1e-12
    assert q(0.975) == pytest.approx(1.959963984540054, abs=1e-8)
    nd = statistics.NormalDist()
    rng = np.random.default_rng(1)
    for u in np.concatenate([rng.uniform(1e-12, 1 - 1e-12, 4000), [1e-12, 0.02424, 0.02426, 0.5, 0.97574, 0.97576]]):
        x, ref = q(float(u)), nd.inv_cdf(float(u))
        assert abs(x - ref) <= 1.2e-9 * abs(ref) + 1e-15, u
    # odd symmetry where 1 - u is exact (dyadic u), and the reference's case
    for k in range(1, 1024):
        assert q(k / 1024) == pytest.approx(-q(1 - k / 1024), rel=1e-12, abs=1e-15)
    assert abs(q(0.99) + q(0.01)) <= 1e-12
    us = np.linspace(1e-12, 1 - 1e-12, 5001)
    xs = [q(float(u)) for u in us]
    assert all(b > a for a, b in zip(xs, xs[1:]))
    assert np.isnan(q(0.0)) and np.isnan(q(1.0))


def args_ptr(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def test_trial_draw_reference_cases(oracle, engine):
    for seed in (0, 7, 2**63 + 5):
        for trial in (1, 2, 99):
            one = E.mc_draw(seed, 5, trial, 1, E.MC_AGAUSS, 1, 0.1, 1)[0]
            many = E.mc_draw(seed, 5, trial, 1, E.MC_AGAUSS, 1, 0.1, 1, 8)[0]
            assert abs(many - 1) >= abs(one - 1)
            assert abs(E.mc_draw(seed, 9, trial, 1, E.MC_LIMIT, 2, 0.5)[0] - 2) == 0.5
            assert abs(E.mc_draw(seed, 9, trial, 1, E.MC_AUNIF, 2, 0.5)[0] - 2) < 0.5
            assert E.mc_draw(seed, 9, trial, 1, E.MC_AGAUSS, 3)[0] == 3


def test_draw_statistics(engine):
    n = 20000
    a = E.mc_draw(7, E.site_key(11, 3), 1, n, E.MC_AGAUSS, 0.7, 0.045, 3)  # sigma 15 mV
    assert abs(a.mean() - 0.7) < 4 * 0.015 / np.sqrt(n)
    assert a.std() == pytest.approx(0.015, rel=0.03)
    g = E.mc_draw(7, 12345, 1, n, E.MC_GAUSS, 1e-7, 0.02, 1)  # 2 % relative
    assert g.std() / 1e-7 == pytest.approx(0.02, rel=0.03)
    u = E.mc_draw(3, 777, 1, n, E.MC_UNIF, 10.0, 0.1)
    assert u.min() >= 9.0 and u.max() <= 11.0
    h, _ = np.histogram(u, bins=10, range=(9, 11))
    assert h.min() > 0.9 * n / 10
    # trials are keyed, not sequenced: a sub-range equals the same trials drawn alone
    assert np.array_equal(a[100:200], E.mc_draw(7, E.site_key(11, 3), 101, 100, E.MC_AGAUSS, 0.7, 0.045, 3))


@pytest.mark.parametrize("fn,args", [
    (0, (1e3, 30.0, 3.0)), (1, (0.7, 0.02, 1.0)), (2, (2.0, 0.5)), (3, (10.0, 0.1)), (4, (2.0, 0.5)),
    (0, (1.0, 0.1, 1.0, 8.0)), (2, (2.0, 0.5, 4.0)),
])
def test_engine_draws_equal_oracle(oracle, engine, fn, args):
    seed, site, n = 0x5eed, E.site_key(42, fn), 512
    got = E.mc_draw(seed, site, 1, n, fn, *args)
    a, ap = args_ptr(args)
    want = np.array([oracle.or_trial_draw(seed, 1 + l, site, fn, ap, len(a)) for l in range(n)])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


# ------------------------------------------------ deck-level Monte Carlo

def agauss_trial_op(oracle, trial, seed=7):
    """One `.dc monte` trial of docs/examples/monte_agauss.sp on the oracle:
    r1v and r2v drawn at their `.param` sites (siteKey(nameKey(nameKey(0,
    ""), name), 0), netlist.zig:3564-3601), then the operating point warm
    from the nominal one (sweep/lanes.zig:73-83). Returns (r1, r2, x)."""
    args = np.array([1000.0, 30.0, 3.0])
    r = [oracle.or_trial_draw(seed, trial, oracle.or_param_site(nm, 0), 0, P(args), 3) for nm in (b"r1v", b"r2v")]
    c = O.Circuit(oracle)
    c.add(O.VSRC, ["in", "0"], {0: 1.0})
    c.add(O.RES, ["in", "out"], {0: 1000.0})
    c.add(O.RES, ["out", "0"], {0: 1000.0})
    c.freeze()
    nom = O.Lanes(c, 1)
    nom.prepare()
    x0 = nom.newton_op(x0=np.zeros(c.n))[0].reshape(-1)
    L = O.Lanes(c, 1)
    L.set_param(O.RES, 0, 0, [r[0]])
    L.set_param(O.RES, 1, 0, [r[1]])
    L.prepare()
    x, st = L.newton_op(x0=x0, init_fix=False)[:2]
    assert st[0] & 1
    return r[0], r[1], np.asarray(x).reshape(-1), c


def test_monte_agauss_rows(oracle):
    """docs/learn/monte-carlo.md:62-69 at full precision: the draws (the
    wyhash site keys, the sampler, Acklam's quantile, agauss's scale)
    reproduce i(v1) and v(in) of the first three trials bit for bit; v(out)
    agrees to within one ulp (trial 3 bit for bit). The reference's last
    bit of v(out) in trials 1-2 is not reproduced by this Newton from the
    nominal point: parity unpinned for that bit."""
    import decks
    for run, i_v1, v_in, v_out in decks.MONTE_AGAUSS_ROWS:
        r1, r2, x, c = agauss_trial_op(oracle, run)
        assert 970 < r1 < 1030 and 970 < r2 < 1030
        assert x[2] == i_v1, run
        assert x[c.node("in")] == v_in, run
        assert abs(x[c.node("out")] - v_out) <= np.spacing(v_out), run
    assert agauss_trial_op(oracle, 3)[2][3] == decks.MONTE_AGAUSS_ROWS[2][3]


def test_wyhash_structure(oracle):
    """The hash's own invariants: seeds and every length class (0, 1-3,
    4-16, 17-48, > 48 bytes) give distinct keys, and a byte change anywhere
    changes the key; the keys that matter are pinned end to end above."""
    seen = set()
    for n in (0, 1, 2, 3, 4, 7, 8, 15, 16, 17, 31, 48, 49, 96, 97, 200):
        key = bytes(range(1, n + 1))
        h = oracle.or_wyhash(0, key, n)
        assert h not in seen
        seen.add(h)
        assert oracle.or_wyhash(1, key, n) != h
        for k in range(n):
            flipped = bytearray(key)
            flipped[k] ^= 0x80
            assert oracle.or_wyhash(0, bytes(flipped), n) != h
