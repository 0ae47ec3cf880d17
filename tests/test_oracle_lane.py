"""The lane bodies of the oracle (src/solver/kernels/lane.zig) against its
scalar LU: a lane's replay of the lead tape is bitwise the scalar refactor
and solve of the same values, failing exactly where the replay fails or a
multiplier passes 1/pivot_tol (lane.zig:811-908)."""
import ctypes as C

import numpy as np

import oracle_lib as O
from test_oracle_lu import random_system


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_lane_replay_is_the_scalar_refactor(oracle):
    rng = np.random.default_rng(0x1A7E51)
    W = 4
    good = failed = 0
    for trial in range(64):
        n = 2 + trial % 20
        cp, ri, v = random_system(rng, n, 0.3)
        s = oracle.or_solver_new(n, O.u32(cp)[1], O.u32(ri)[1])
        if oracle.or_solver_factor(s, O.f64(v)[1]) != 0 or oracle.or_is_tridiag(n, O.u32(cp)[1], O.u32(ri)[1]):
            oracle.or_solver_free(s)
            continue
        lu = oracle.or_solver_lu(s)
        tab = O.LaneTab()
        ln = C.c_size_t()
        idxp = oracle.or_build_tab(lu, O.u32(cp)[1], oracle.or_solver_q(s), 1e-12, 1e-3, C.byref(tab), C.byref(ln))
        idx = np.ctypeslib.as_array(idxp, (ln.value,)).copy()
        oracle.free_(C.cast(idxp, C.c_void_p))
        nnz = len(v)
        a = np.repeat(v[:, None], W, axis=1)
        for l in range(W):
            k = rng.integers(6)
            if k == 0:
                a[rng.integers(nnz), l] = 4 * rng.random() - 2
            elif k == 1:
                c = rng.integers(n)
                a[cp[c]:cp[c + 1], l] = 0
            else:
                a[:, l] = v * (0.95 + 0.1 * rng.random(nnz))
        rhs = 2 * rng.random((n, W)) - 1
        lx = np.zeros((max(tab.nnz_l, 1), W))
        ux = np.zeros((max(tab.nnz_u, 1), W))
        ud, w, y, dx = (np.zeros((n, W)) for _ in range(4))
        a = np.ascontiguousarray(a)
        for l in range(W):
            bad = oracle.or_lane_refactor(C.byref(tab), O.u32(idx)[1], dp(a), dp(lx), dp(ux), dp(ud), dp(w), W, l)
            assert not np.any(w), "w must be zero on exit"
            # the scalar oracle: the same pivots replayed, then the multipliers
            col = np.ascontiguousarray(a[:, l])
            r = oracle.or_lu_refactor(lu, O.u32(cp)[1], O.f64(col)[1], 1e-12)
            wide = False
            if r == 0:
                nl = oracle.or_lu_nnz_l(lu)
                sp = np.ctypeslib.as_array(oracle.or_lu_scaled_pivot(lu), (n,))
                lx_s = np.ctypeslib.as_array(oracle.or_lu_lx(lu), (nl,)) if nl else np.zeros(0)
                lp = idx[tab.lp:tab.lp + n + 1]
                worst = max([abs(lx_s[p]) for k in range(n) if not sp[k] for p in range(lp[k], lp[k + 1])] or [0.0])
                wide = not (worst <= 1e3)
            expect_bad = r != 0 or wide
            assert bool(bad) == expect_bad
            if bad:
                failed += 1
                continue
            good += 1
            oracle.or_lane_solve_neg(C.byref(tab), O.u32(idx)[1], dp(lx), dp(ux), dp(ud), dp(np.ascontiguousarray(rhs)),
                                     dp(y), dp(dx), W, l)
            ref, refp = O.f64(np.zeros(n))
            oracle.or_lu_solve(lu, O.f64(-rhs[:, l])[1], refp)
            assert np.array_equal(dx[:, l].view(np.uint64), ref.view(np.uint64))
        oracle.or_solver_free(s)
    assert good > 0 and failed > 0
