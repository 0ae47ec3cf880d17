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

GAMMA = 0x9e3779b97f4a7c15
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
