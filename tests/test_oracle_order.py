"""Column ordering (BTF + AMD + Cholesky count): the oracle pinned to the
reference's known answers (src/solver/tests.zig:2610-2781) and properties
(src/solver/order.zig:721-785), the Cholesky count to a brute-force
symbolic elimination, and the engine's host ordering equal to the oracle's
on every pattern."""
import itertools

import numpy as np

import oracle_lib as O


def dense_csc(a):
    a = np.asarray(a)
    n = a.shape[0]
    col_ptr, row_idx = [0], []
    for j in range(n):
        for i in range(n):
            if a[i, j] != 0:
                row_idx.append(i)
        col_ptr.append(len(row_idx))
    return n, np.array(col_ptr, np.uint32), np.array(row_idx, np.uint32)


def order(lib, a, fn="or_order"):
    n, cp, ri = a if isinstance(a, tuple) else dense_csc(a)
    q, qp = O.u32(np.zeros(n))
    assert getattr(lib, fn)(n, O.u32(cp)[1], O.u32(ri)[1], qp) == 0
    return q


def random_pattern(rng, n, density):
    col_ptr, row_idx = [0], []
    for j in range(n):
        for i in range(n):
            if i == j or rng.random() < density:
                row_idx.append(i)
        col_ptr.append(len(row_idx))
    return n, np.array(col_ptr, np.uint32), np.array(row_idx, np.uint32)


STAR = [[1, 1, 1, 1, 1], [1, 1, 0, 0, 0], [1, 0, 1, 0, 0], [1, 0, 0, 1, 0], [1, 0, 0, 0, 1]]


def test_star_defers_the_hub(oracle):
    for fn in ("or_order", "or_amd"):
        q = order(oracle, STAR, fn)
        assert sorted(q) == list(range(5))
        assert q[3] == 0 or q[4] == 0


def test_btf_sink_block_first(oracle):
    q = order(oracle, [[1, 1, 1, 0], [1, 1, 0, 0], [0, 0, 1, 1], [0, 0, 1, 1]])
    assert q[0] < 2 and q[1] < 2 and q[2] >= 2 and q[3] >= 2


def test_btf_three_blocks_in_topological_order(oracle):
    assert list(order(oracle, [[1, 1, 1], [0, 1, 1], [0, 0, 1]])) == [0, 1, 2]


def test_arrowhead_hub_last(oracle):
    n = 8
    a = np.eye(n)
    a[0, :] = 1
    a[:, 0] = 1
    q = order(oracle, a, "or_amd")
    assert q[-1] == 0 or q[-2] == 0


def test_chain_complete_graph_identity_and_1x1(oracle):
    a = np.eye(6) + np.eye(6, k=1) + np.eye(6, k=-1)
    assert sorted(order(oracle, a, "or_amd")) == list(range(6))
    assert sorted(order(oracle, np.ones((4, 4)), "or_amd")) == list(range(4))
    assert sorted(order(oracle, np.eye(3))) == [0, 1, 2]
    assert list(order(oracle, np.eye(1))) == [0]


def test_random_patterns_are_permutations(oracle):
    rng = np.random.default_rng(0x0DE7)
    for t in range(300):
        n = t + 1 if t < 64 else int(rng.integers(1, 65))
        density = 0.3 + 0.6 * rng.random() if t % 4 == 0 else 0.05 * rng.random()
        pat = random_pattern(rng, n, density)
        for fn in ("or_order", "or_amd"):
            q = order(oracle, pat, fn)
            assert sorted(q) == list(range(n))


def brute_chol_count(n, cp, ri, perm):
    """Entries of the Cholesky factor of A + A^T (diagonal included) under
    `perm`, by explicit symbolic elimination on sets."""
    pinv = np.empty(n, int)
    pinv[np.asarray(perm)] = np.arange(n)
    adj = [set() for _ in range(n)]
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            a, b = pinv[ri[p]], pinv[j]
            if a != b:
                adj[a].add(b)
                adj[b].add(a)
    count = 0
    for k in range(n):
        later = {v for v in adj[k] if v > k}
        count += 1 + len(later)
        for u, v in itertools.combinations(later, 2):
            adj[u].add(v)
            adj[v].add(u)
    return count


def test_chol_count_is_the_symbolic_elimination(oracle):
    rng = np.random.default_rng(0x0DE9)
    for _ in range(60):
        n = int(rng.integers(1, 30))
        pat = random_pattern(rng, n, 0.15)
        perm = rng.permutation(n).astype(np.uint32)
        got = oracle.or_chol_count(n, O.u32(pat[1])[1], O.u32(pat[2])[1], O.u32(perm)[1], 2 ** 64 - 1)
        assert got == brute_chol_count(*pat, perm)


def test_chol_count_cap(oracle):
    rng = np.random.default_rng(0x0DEA)
    n = 40
    for _ in range(30):
        pat = random_pattern(rng, n, 0.08)
        perm = rng.permutation(n).astype(np.uint32)
        full = oracle.or_chol_count(n, O.u32(pat[1])[1], O.u32(pat[2])[1], O.u32(perm)[1], 2 ** 64 - 1)
        for cap in (n, full - 1, full, full + 1):
            got = oracle.or_chol_count(n, O.u32(pat[1])[1], O.u32(pat[2])[1], O.u32(perm)[1], cap)
            assert got == full if full < cap else got >= cap


def test_order_never_fills_more_than_natural_where_it_chose(oracle):
    """order.zig:133-146: each block keeps AMD or natural, whichever the
    Cholesky count prefers, so a single-block pattern never fills more than
    its natural order (tests.zig:2799, #46)."""
    rng = np.random.default_rng(42)
    for _ in range(40):
        n = int(rng.integers(4, 40))
        pat = random_pattern(rng, n, 0.2)
        q = order(oracle, pat)
        nat = np.arange(n, dtype=np.uint32)
        c = lambda p: oracle.or_chol_count(n, O.u32(pat[1])[1], O.u32(pat[2])[1], O.u32(p)[1], 2 ** 64 - 1)
        blocks_emit, bp = O.u32(np.zeros(n)), O.u32(np.zeros(n + 1))
        nb = oracle.or_blocks(n, O.u32(pat[1])[1], O.u32(pat[2])[1], blocks_emit[1], bp[1])
        if nb == 1:
            assert c(q) <= c(nat)


def test_engine_ordering_equals_oracle(oracle, hostlane):
    rng = np.random.default_rng(7)
    for t in range(200):
        n = int(rng.integers(1, 80))
        pat = random_pattern(rng, n, rng.choice([0.02, 0.05, 0.2, 0.6]))
        qo = order(oracle, pat)
        qe, qp = O.u32(np.zeros(n))
        assert hostlane.hl_column_order(n, O.u32(pat[1])[1], O.u32(pat[2])[1], qp) == 0
        assert np.array_equal(qo, qe), f"trial {t} n={n}"
        ao = order(oracle, pat, "or_amd")
        ae, ap = O.u32(np.zeros(n))
        assert hostlane.hl_amd(n, O.u32(pat[1])[1], O.u32(pat[2])[1], ap) == 0
        assert np.array_equal(ao, ae)
