"""Gilbert-Peierls LU and the direct-solver facade of the oracle, pinned to
the reference's own LU tests (src/solver/sparse_lu.zig:1450-1740) and to
LAPACK's scaled residual, plus the dense peel against numpy."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O


def random_system(rng, n, density):
    """sparse_lu.zig:1380-1410 TestSystem."""
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


class LU:
    def __init__(self, lib, n, cp, q):
        self.lib, self.n = lib, n
        self.cp, self.cpp = O.u32(cp)
        self.q, self.qp = O.u32(q)
        self.h = lib.or_lu_new(n, self.cpp, self.qp)

    def __del__(self):
        self.lib.or_lu_free(self.h)

    def factor(self, ri, vals, tol=1e-3):
        return self.lib.or_lu_factor(self.h, self.cpp, O.u32(ri)[1], O.f64(vals)[1], tol)

    def refactor(self, vals, growth=1e-12):
        return self.lib.or_lu_refactor(self.h, self.cpp, O.f64(vals)[1], growth)

    def solve(self, b, transpose=False):
        x, xp = O.f64(np.zeros(self.n))
        (self.lib.or_lu_solve_t if transpose else self.lib.or_lu_solve)(self.h, O.f64(b)[1], xp)
        return x

    def arrays(self):
        nl, nu = self.lib.or_lu_nnz_l(self.h), self.lib.or_lu_nnz_u(self.h)
        def get(f, k):
            if k == 0:
                return np.zeros(0, dtype=np.float64)
            return np.ctypeslib.as_array(f(self.h), (k,)).copy()
        return (get(self.lib.or_lu_lx, nl), get(self.lib.or_lu_ux, nu), get(self.lib.or_lu_udiag, self.n),
                get(self.lib.or_lu_pinv, self.n))

    def w_zero(self):
        return not np.any(np.ctypeslib.as_array(self.lib.or_lu_w(self.h), (max(self.n, 1),))[:self.n])


def scaled_residual(cp, ri, vals, x, b, transpose=False):
    """LAPACK ||A x - b|| / (n max|a| ||x|| n eps), max norms
    (sparse_lu.zig:1419-1439)."""
    n = len(cp) - 1
    r = np.array(b, dtype=float)
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            i = ri[p]
            if transpose:
                r[j] -= vals[p] * x[i]
            else:
                r[i] -= vals[p] * x[j]
    return np.abs(r).max() / (n * np.abs(vals).max() * np.abs(x).max() * n * np.finfo(float).eps)


def test_factor_refactor_solve_random(oracle):
    rng = np.random.default_rng(0x5BA75E)
    solved = 0
    for t in range(256):
        n = t + 1 if t < 64 else int(rng.integers(1, 65))
        cp, ri, v = random_system(rng, n, rng.random() * (0.6 if n > 32 else 0.3))
        q = rng.permutation(n).astype(np.uint32)
        lu = LU(oracle, n, cp, q)
        if lu.factor(ri, v) != 0:
            assert lu.w_zero()
            continue
        assert lu.w_zero()
        solved += 1
        b = 2 * rng.random(n) - 1
        for tr in (False, True):
            x = lu.solve(b, tr)
            assert scaled_residual(cp, ri, v, x, b, tr) <= 1e6
        f_l, f_u, f_d, _ = lu.arrays()
        # the full factor is bitwise a column replay of the same values
        assert lu.refactor(v) == 0
        r_l, r_u, r_d, _ = lu.arrays()
        assert np.array_equal(f_l.view(np.uint64), r_l.view(np.uint64))
        assert np.array_equal(f_u.view(np.uint64), r_u.view(np.uint64))
        assert np.array_equal(f_d.view(np.uint64), r_d.view(np.uint64))
        pert = v * (0.9 + 0.2 * rng.random(len(v)))
        if lu.refactor(pert) == 0:
            assert lu.w_zero()
            x = lu.solve(b)
            assert scaled_residual(cp, ri, pert, x, b) <= 1e6
        assert lu.w_zero()
    assert solved >= 128


def test_void_column_nan_replay_leaves_w_zero(oracle):
    cp = np.array([0, 2, 4, 5, 6], np.uint32)
    ri = np.array([0, 3, 0, 1, 2, 3], np.uint32)
    lu = LU(oracle, 4, cp, np.arange(4, dtype=np.uint32))
    assert lu.factor(ri, [1, 0.5, 0, 0, 1, 1]) == 0
    vc = np.ctypeslib.as_array(oracle.or_lu_void_col(lu.h), (4,))
    assert vc[1] == 1
    assert lu.refactor([1, np.nan, 0, 0, 0, 1]) == 1
    assert lu.w_zero()


def test_overflow_off_the_pivot_fails(oracle):
    cp = np.array([0, 2, 3, 6], np.uint32)
    ri = np.array([0, 1, 1, 0, 1, 2], np.uint32)
    benign, over = [1, 1, 1, 1, 1, 1], [1, 1, 1, 1.5e308, -1.5e308, 1]
    lu = LU(oracle, 3, cp, np.arange(3, dtype=np.uint32))
    assert lu.factor(ri, over) == 1 and lu.w_zero()
    assert lu.factor(ri, benign) == 0
    assert lu.refactor(over) == 1 and lu.w_zero()
    assert lu.refactor(over, 0.0) == 1
    assert lu.refactor(benign) == 0


def test_fill_cap_restarts_on_the_diagonal(oracle):
    n = 6
    cp, ri, v = [0], [], []
    for j in range(n):
        for i in range(n):
            if i == j or i == n - 1 or j == n - 1:
                ri.append(i)
                v.append((1e-6 if i == 0 else 4) if i == j else 1)
        cp.append(len(ri))
    for diag_fill, ok in ((10, True), (9, False)):
        lu = LU(oracle, n, cp, np.arange(n, dtype=np.uint32))
        oracle.or_lu_set_caps(lu.h, 1.0, 1e-12, diag_fill)
        r = lu.factor(ri, v, 0.5)
        assert lu.w_zero()
        if not ok:
            assert r == 1
            continue
        assert r == 0
        _, _, _, pinv = lu.arrays()
        assert list(pinv) == list(range(n))
        assert oracle.or_lu_nnz_l(lu.h) + oracle.or_lu_nnz_u(lu.h) == 10


def test_empty_and_one_by_one(oracle):
    lu = LU(oracle, 1, [0, 1], [0])
    assert lu.factor([0], [4.0]) == 0
    assert lu.solve([2.0])[0] == 0.5
    assert lu.refactor([np.inf]) == 1
    assert lu.factor([0], [np.nan]) == 1 and lu.w_zero()
    empty = LU(oracle, 0, [0], [])
    assert empty.factor([], []) == 0 and empty.refactor([]) == 0


def test_power_of_two_scaling_is_exact(oracle):
    rng = np.random.default_rng(0x2B17)
    sa, sb, sx = 2.0 ** -20, 2.0 ** 7, 2.0 ** 27
    checked = 0
    for t in range(64):
        n = 1 + t % 48
        cp, ri, v = random_system(rng, n, 0.5 if n > 32 else 0.25)
        q = rng.permutation(n).astype(np.uint32)
        b = 2 * rng.random(n) - 1
        a, a2 = LU(oracle, n, cp, q), LU(oracle, n, cp, q)
        ok, ok2 = a.factor(ri, v) == 0, a2.factor(ri, v * sa) == 0
        assert ok == ok2
        if not ok:
            continue
        assert np.array_equal(a.arrays()[3], a2.arrays()[3])
        for tr in (False, True):
            assert np.array_equal((a.solve(b, tr) * sx).view(np.uint64), a2.solve(b * sb, tr).view(np.uint64))
        checked += 1
    assert checked >= 32


def test_solver_bypass_replay_and_repivot(oracle):
    """direct.Solver: unchanged values bypass, a replay keeps its pivots,
    and a replay that fails re-pivots with a full factor."""
    rng = np.random.default_rng(11)
    cp, ri, v = random_system(rng, 12, 0.3)
    s = oracle.or_solver_new(12, O.u32(cp)[1], O.u32(ri)[1])
    assert oracle.or_solver_factor(s, O.f64(v)[1]) == 0
    b = 2 * rng.random(12) - 1
    x1, xp1 = O.f64(np.zeros(12))
    oracle.or_solver_solve_neg(s, O.f64(b)[1], xp1)
    assert scaled_residual(cp, ri, v, -x1, b) <= 1e6
    assert oracle.or_solver_factor(s, O.f64(v)[1]) == 0  # bypass
    v2 = v.copy()
    for j in range(12):  # kill the chosen pivots: the replay fails, re-pivot
        d = [p for p in range(cp[j], cp[j + 1]) if ri[p] == j]
        if d:
            v2[d[0]] = 0.0
    r = oracle.or_solver_factor(s, O.f64(v2)[1])
    if r == 0:
        x2, xp2 = O.f64(np.zeros(12))
        oracle.or_solver_solve_neg(s, O.f64(b)[1], xp2)
        assert scaled_residual(cp, ri, v2, -x2, b) <= 1e6
    oracle.or_solver_free(s)


def test_dense_peel_matches_numpy(oracle):
    rng = np.random.default_rng(5)
    for n in range(1, 40):
        m = rng.standard_normal((n, n)) + n * np.eye(n)
        b = rng.standard_normal(n)
        mm, mp = O.f64(m.reshape(-1).copy())
        x, xp = O.f64(b.copy())
        assert oracle.or_dense_solve_neg(n, mp, xp) == 1
        assert np.allclose(x, -np.linalg.solve(m, b), rtol=1e-10, atol=1e-12)
    sing, sp = O.f64(np.zeros(4))
    xs, xsp = O.f64(np.ones(2))
    assert oracle.or_dense_solve_neg(2, sp, xsp) == 0
