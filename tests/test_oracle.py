"""Pins the CPU oracle against routes that do not share its code.

Each test names what it pins:
  * scipy.sparse.csgraph shortest paths (Dijkstra / unweighted BFS),
  * numpy.linalg.solve for the PageRank fixed point (closed form),
  * pure-Python brute force (Bellman-Ford, reachability, list edits),
  * published known answers (splitmix64) and hand-derived fixtures.
A planted mistake in a result vector must be caught by the restated --check
functions (hybrid_sssp.cu:211-316, hybrid_cc_host.cu:8-32, hybrid_pr_host.cu:9-56).
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import breadth_first_order, dijkstra, shortest_path

from oracle import oracle as O

INF = O.INF


def random_multigraph(rng, n, m, dup_frac=0.1, self_loops=True):
    src = rng.integers(0, n, m, dtype=np.int64)
    dst = rng.integers(0, n, m, dtype=np.int64)
    k = int(m * dup_frac)
    if k:
        idx = rng.integers(0, m, k)
        src = np.concatenate([src, src[idx]])
        dst = np.concatenate([dst, dst[idx]])
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    return src.astype(np.uint32), dst.astype(np.uint32)


def weights_py(mode, seed, src, dst):
    return np.array([O.weight(mode, seed, int(u), int(v)) for u, v in zip(src, dst)],
                    dtype=np.float64)


def min_weight_csr(n, src, dst, w):
    """CSR with the minimum weight over parallel edges (scipy sums duplicates)."""
    order = np.lexsort((w, dst, src))
    s, d, ww = src[order], dst[order], w[order]
    first = np.ones(len(s), dtype=bool)
    first[1:] = (s[1:] != s[:-1]) | (d[1:] != d[:-1])
    return sp.csr_matrix((ww[first], (s[first], d[first])), shape=(n, n))


def to_u32(dist):
    out = np.full(dist.shape, INF, dtype=np.uint64)
    fin = np.isfinite(dist)
    out[fin] = dist[fin].astype(np.uint64)
    return out.astype(np.uint32)


def test_splitmix64_known_answers():
    # Published first two outputs of SplitMix64 seeded with 0 (state advances
    # by the golden gamma, output = mix(state)).
    assert O.lib().og_splitmix64(0) == 0xE220A8397B1DCDAF
    assert O.lib().og_splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_weight_modes_closed_form():
    assert O.weight(O.W_REF128, 0, 0, 0) == 1
    assert O.weight(O.W_REF128, 0, 127, 1) == 1
    assert O.weight(O.W_REF128, 0, 63, 1) == 65
    # 64-bit sum: no 32-bit wrap for ids near 2^32
    assert O.weight(O.W_REF128, 0, 0xFFFFFFFF, 1) == (0xFFFFFFFF + 1) % 128 + 1
    ws = [O.weight(O.W_HASH64, 12345, u, v) for u in range(50) for v in range(50)]
    assert min(ws) == 1 and max(ws) == 64
    assert O.weight(O.W_UNIT, 0, 3, 4) == 1 and O.weight(O.W_ZERO, 0, 3, 4) == 0


@pytest.mark.parametrize("mode,seed", [(O.W_REF128, 0), (O.W_HASH64, 7)])
@pytest.mark.parametrize("n,m,case", [(60, 300, 0), (500, 4000, 1), (2000, 6000, 2)])
def test_dijkstra_vs_scipy(mode, seed, n, m, case):
    rng = np.random.default_rng(100 + case)
    src, dst = random_multigraph(rng, n, m)
    g = O.Graph.from_edges(n, src, dst)
    w = weights_py(mode, seed, src, dst)
    ref = dijkstra(min_weight_csr(n, src, dst, w), directed=True, indices=0)
    got = g.sssp(0, mode, seed)
    assert np.array_equal(got, to_u32(ref))
    assert np.array_equal(g.sssp_label_correcting(0, mode, seed), got)


def test_dijkstra_vs_bellman_ford_bruteforce():
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = int(rng.integers(1, 12))
        m = int(rng.integers(0, 30))
        src, dst = random_multigraph(rng, n, m)
        g = O.Graph.from_edges(n, src, dst)
        s = int(rng.integers(0, n))
        d = [None] * n
        d[s] = 0
        for _ in range(n):
            for u, v in zip(src.tolist(), dst.tolist()):
                if d[u] is not None:
                    c = d[u] + (u + v) % 128 + 1
                    if d[v] is None or c < d[v]:
                        d[v] = c
        exp = np.array([INF if x is None else x for x in d], dtype=np.uint32)
        assert np.array_equal(g.sssp(s, O.W_REF128), exp), trial


@pytest.mark.parametrize("n,m", [(100, 300), (3000, 9000)])
def test_bfs_vs_scipy(n, m):
    rng = np.random.default_rng(n)
    src, dst = random_multigraph(rng, n, m)
    g = O.Graph.from_edges(n, src, dst)
    A = sp.csr_matrix((np.ones(len(src)), (src, dst)), shape=(n, n))
    ref = shortest_path(A, directed=True, unweighted=True, indices=3)
    assert np.array_equal(g.bfs(3), to_u32(ref))


def test_cc_flood_vs_reachability_bruteforce():
    rng = np.random.default_rng(11)
    for trial in range(10):
        n = int(rng.integers(1, 60))
        m = int(rng.integers(0, 3 * n))
        src, dst = random_multigraph(rng, n, m)
        g = O.Graph.from_edges(n, src, dst)
        A = sp.csr_matrix((np.ones(len(src)), (src, dst)), shape=(n, n))
        exp = np.full(n, INF, dtype=np.uint32)
        for u in range(n):  # label[v] = min{u : u reaches v}
            reach = breadth_first_order(A, u, directed=True, return_predecessors=False)
            for v in reach:
                exp[v] = min(exp[v], u)
        got = g.cc()
        assert np.array_equal(got, exp), trial
        assert g.cc_check(got) == 0
        if n > 1:
            bad = got.copy()
            bad[n - 1] ^= 1
            assert g.cc_check(bad) >= 1  # planted mistake is caught


@pytest.mark.parametrize("n,m", [(40, 120), (300, 1500)])
def test_pagerank_vs_linear_solve(n, m):
    rng = np.random.default_rng(n + m)
    src, dst = random_multigraph(rng, n, m)
    g = O.Graph.from_edges(n, src, dst)
    deg = np.bincount(src, minlength=n).astype(np.float64)
    M = np.zeros((n, n))
    for u, v in zip(src, dst):  # parallel edges count; dangling rows stay zero
        M[v, u] += 0.85 / deg[u]
    exact = np.linalg.solve(np.eye(n) - M, np.full(n, 0.15))
    x, rounds = g.pagerank()
    assert rounds > 0
    assert np.max(np.abs(x - exact)) <= 1e-9 * max(1.0, np.max(exact))
    # The restated PageRankCheck accepts the converged ranks with the exact
    # float residual and rejects a planted rank error.
    ranks = exact.astype(np.float32)
    nxt = 0.15 + M @ ranks.astype(np.float64)
    residual = (nxt - ranks).astype(np.float32)
    res = g.pr_check(ranks, residual, 1e-5)
    assert res.ok and res.converged and res.finite
    bad = ranks.copy()
    bad[0] += 1.0
    assert not g.pr_check(bad, residual, 1e-5).ok


def brute_apply(rows, dels, adds, n):
    """List edits with the reference semantics, written independently."""
    rows = [list(r) for r in rows]
    miss = inval = 0
    from collections import Counter, defaultdict
    req = defaultdict(Counter)
    for u, v in dels:
        if u >= n or v >= n:
            inval += 1
            continue
        req[u][v] += 1
    for u, c in req.items():
        out = []
        left = dict(c)
        for d in rows[u]:
            if left.get(d, 0) > 0:
                left[d] -= 1
            else:
                out.append(d)
        miss += sum(left.values())
        rows[u] = out
    for u, v in adds:
        if u >= n or v >= n:
            inval += 1
            continue
        rows[u].append(v)
    return rows, miss, inval


def test_mutation_semantics_vs_bruteforce():
    rng = np.random.default_rng(3)
    for trial in range(30):
        n = int(rng.integers(1, 30))
        m = int(rng.integers(0, 120))
        src, dst = random_multigraph(rng, n, m, dup_frac=0.3)
        g = O.Graph.from_edges(n, src, dst)
        rows = g.rows()
        # deletes: mostly present edges (some duplicated), some missing, some invalid
        k = int(rng.integers(0, 40))
        dels = []
        for _ in range(k):
            r = rng.random()
            if r < 0.7 and m:
                i = int(rng.integers(0, len(src)))
                dels.append((int(src[i]), int(dst[i])))
            elif r < 0.9:
                dels.append((int(rng.integers(0, n)), int(rng.integers(0, n))))
            else:
                dels.append((int(rng.integers(0, n + 3)), n + 1))
        adds = [(int(rng.integers(0, n + 1)), int(rng.integers(0, n + 1))) for _ in range(k)]
        exp_rows, exp_miss, exp_inval = brute_apply(rows, dels, adds, n)
        du = np.array([d[0] for d in dels], dtype=np.uint32)
        dv = np.array([d[1] for d in dels], dtype=np.uint32)
        miss, inval = g.apply_deletes(du, dv)
        au = np.array([a[0] for a in adds], dtype=np.uint32)
        av = np.array([a[1] for a in adds], dtype=np.uint32)
        inval += g.apply_adds(au, av)
        assert g.rows() == exp_rows, trial  # order-exact, not only multiset
        assert (miss, inval) == (exp_miss, exp_inval)
        # the worker-pool form (any thread count) gives the same rows, in order
        for threads in (1, 3, 8):
            gm = O.Graph.from_edges(n, src, dst)
            assert gm.apply_batch_mt(du, dv, au, av, threads) == (exp_miss, exp_inval), (trial, threads)
            assert gm.rows() == exp_rows, (trial, threads)


def test_bellman_and_parent_checks_catch_planted_mistakes():
    rng = np.random.default_rng(9)
    n, m = 400, 3000
    src, dst = random_multigraph(rng, n, m, self_loops=False)
    g = O.Graph.from_edges(n, src, dst)
    dist = g.sssp(0)
    res = g.check_bellman(dist, 0)
    assert res.relaxable_edges == 0 and res.missing_tight_witnesses == 0
    # parents: smallest tight in-neighbour, by brute force
    parent = np.full(n, INF, dtype=np.uint32)
    for u, v in zip(src.tolist(), dst.tolist()):
        if dist[u] != INF and dist[v] != INF and v != 0 and \
                int(dist[u]) + (u + v) % 128 + 1 == int(dist[v]):
            parent[v] = min(parent[v], u)
    pres = g.check_parent(dist, parent, 0)
    assert pres.invalid_vertices == 0
    reach = np.nonzero((dist != INF) & (np.arange(n) != 0))[0]
    v = int(reach[len(reach) // 2])
    low = dist.copy()
    low[v] -= 1  # no tight witness any more
    assert g.check_bellman(low, 0).missing_tight_witnesses >= 1
    high = dist.copy()
    high[v] += 1  # its tight in-edge is now relaxable
    assert g.check_bellman(high, 0).relaxable_edges >= 1
    wrong = parent.copy()
    wrong[v] = (int(parent[v]) + 1) % n
    assert g.check_parent(dist, wrong, 0).invalid_vertices >= 1


def test_checksums_special_cases():
    basis = 1469598103934665603
    assert O.distance_checksum(np.array([], dtype=np.uint32)) == basis
    assert O.distance_checksum(np.array([INF, INF], dtype=np.uint32)) == basis
    # unreachable entries are skipped; order matters
    a = O.distance_checksum(np.array([0, 5, INF], dtype=np.uint32))
    b = O.distance_checksum(np.array([0, 5], dtype=np.uint32))
    c = O.distance_checksum(np.array([5, 0], dtype=np.uint32))
    assert a == b and a != c
    # single entry: (basis ^ (0 + golden + 0 + 0)) * prime mod 2^64
    one = O.distance_checksum(np.array([0], dtype=np.uint32))
    assert one == ((basis ^ 0x9E3779B97F4A7C15) * 1099511628211) % (1 << 64)


def test_parent_checksum_closed_forms():
    """og_parent_checksum (hybrid_sssp.cu:604-619): basis 1099511628211,
    prime 1469598103934665603, golden 0x517cc1b727220a95, parent terms
    (p<<7) + (p>>3); unreachable vertices are skipped."""
    basis, prime, M = 1099511628211, 1469598103934665603, 1 << 64
    e = np.array([], dtype=np.uint32)
    assert O.parent_checksum(e, e) == basis
    inf2 = np.array([INF, INF], dtype=np.uint32)
    assert O.parent_checksum(inf2, np.array([3, 9], dtype=np.uint32)) == basis

    def one(v, p):
        return ((basis ^ (v + 0x517CC1B727220A95 + (p << 7) + (p >> 3))) * prime) % M

    # a single reachable vertex at index v: the INF entries before it are skipped
    for v, p in ((0, 0), (0, 12345), (2, 4000000000), (1, INF)):
        dist = np.full(v + 1, INF, dtype=np.uint32)
        dist[v] = 7
        par = np.zeros(v + 1, dtype=np.uint32)
        par[v] = p
        assert O.parent_checksum(dist, par) == one(v, p), (v, p)
    # two entries chain: the second folds into the first's product
    d = np.array([0, 4], dtype=np.uint32)
    pr = np.array([INF, 0], dtype=np.uint32)
    h = one(0, INF)
    h = ((h ^ (1 + 0x517CC1B727220A95 + 0 + 0)) * prime) % M
    assert O.parent_checksum(d, pr) == h
    # order and the parent value matter; the distance value does not
    assert O.parent_checksum(d, pr) != O.parent_checksum(d, pr[::-1].copy())
    assert O.parent_checksum(d, pr) == O.parent_checksum(np.array([9, 1], dtype=np.uint32), pr)
    pr2 = pr.copy()
    pr2[1] = 1
    assert O.parent_checksum(d, pr2) != O.parent_checksum(d, pr)


def test_pr_error_helper_hand_vectors():
    """The normalised L1 / max error used by every PageRank parity test:
    sum|x - x*| / sum|x*| and max|x - x*|, on hand-computed vectors."""
    from tests.helpers import pr_errors
    ref = np.array([1.0, 2.0, -1.0, 0.5])
    x = np.array([1.5, 2.0, -0.75, 0.5], dtype=np.float32)
    l1n, mx = pr_errors(x, ref)
    assert abs(l1n - 0.75 / 4.5) < 1e-15
    assert mx == 0.5
    l1n, mx = pr_errors(ref.astype(np.float32), ref)
    assert l1n == 0.0 and mx == 0.0


@pytest.mark.parametrize("n,m", [(40, 120), (300, 1500)])
def test_pr_invariant_closed_forms(n, m):
    """og_pr_invariant (PageRankCheck's invariant, hybrid_pr_host.cu:15-45):
    x = 0 with r = 0.15 has zero defect; the exact solve (numpy) with r = 0
    has ~zero defect; moving x[v] by delta adds delta * (1 + 0.85 * (non-self
    out-edges of v) / deg(v)) for a vertex v with out-edges, delta alone for a
    dangling one; a non-finite entry is flagged."""
    rng = np.random.default_rng(7 * n + m)
    src, dst = random_multigraph(rng, n, m)
    g = O.Graph.from_edges(n, src, dst)
    deg = np.bincount(src, minlength=n).astype(np.float64)
    M = np.zeros((n, n))
    for u, v in zip(src, dst):
        M[v, u] += 0.85 / deg[u]
    zero = np.zeros(n, dtype=np.float32)
    ok, d, mr, mass = g.pr_invariant(zero, np.full(n, 0.15, dtype=np.float32))
    # only the float rounding of 0.15 is left
    assert ok and abs(d - n * abs(0.15 - float(np.float32(0.15)))) < 1e-12
    assert mr == float(np.float32(0.15)) and mass == 0.0
    exact = np.linalg.solve(np.eye(n) - M, np.full(n, 0.15))
    x = exact.astype(np.float32)
    r = (0.15 + M @ x.astype(np.float64) - x.astype(np.float64)).astype(np.float32)
    ok, d0, _, mass = g.pr_invariant(x, r)
    assert ok and d0 <= 1e-5 * n and abs(mass - np.abs(x.astype(np.float64)).sum()) < 1e-9
    delta = 0.5
    for v in (int(np.argmax(deg)), int(np.argmin(deg))):
        xb = x.copy()
        xb[v] += np.float32(delta)
        _, d1, _, _ = g.pr_invariant(xb, r)
        if deg[v] > 0:
            loops = int(np.sum((src == v) & (dst == v)))
            # x[v] itself moves by delta; 0.85*delta/deg(v) lands on each out-edge's
            # destination, v included for each self-loop
            exp = abs(delta * (0.85 * loops / deg[v] - 1.0)) + delta * 0.85 * (deg[v] - loops) / deg[v]
        else:
            exp = delta
        assert abs((d1 - d0) - exp) <= 1e-4, (v, d1 - d0, exp)
    nanx = x.copy()
    nanx[0] = np.nan
    assert not g.pr_invariant(nanx, r)[0]


def test_check_parent_brute_force():
    """og_check_parent (CheckParentWitness, hybrid_sssp.cu:266-316) against a
    direct restatement over Python sets: for every reached v other than the
    source, its parent p must be reached, hold v in its row and satisfy
    dist[p] + w(p,v) == dist[v]; parents are drawn from true tight parents,
    random vertices, INF and out-of-range ids."""
    rng = np.random.default_rng(31)
    for trial in range(30):
        n = int(rng.integers(2, 50))
        src, dst = random_multigraph(rng, n, int(rng.integers(1, 4 * n)))
        g = O.Graph.from_edges(n, src, dst)
        dist = g.sssp(0, O.W_REF128)
        edges = set(zip(src.tolist(), dst.tolist()))
        par = rng.integers(0, n, n).astype(np.uint32)
        pick = rng.random(n)
        par[pick < 0.1] = INF
        par[(pick >= 0.1) & (pick < 0.15)] = n + 3
        for v in range(n):  # most vertices get a true tight parent
            if pick[v] >= 0.5 and dist[v] != INF:
                for u, w in edges:
                    if w == v and dist[u] != INF and int(dist[u]) + O.weight(O.W_REF128, 0, u, v) == dist[v]:
                        par[v] = u
                        break
        inval = missing = wrong = 0
        first = INF
        for v in range(n):
            if v == 0 or dist[v] == INF:
                continue
            p = int(par[v])
            ok = p < n and dist[p] != INF
            if ok:
                exists = (p, v) in edges
                matches = int(dist[p]) + O.weight(O.W_REF128, 0, p, v) == int(dist[v])
                missing += not exists
                wrong += exists and not matches
                ok = exists and matches
            if not ok:
                first = v if inval == 0 else first
                inval += 1
        r = g.check_parent(dist, par, 0, O.W_REF128)
        assert (r.invalid_vertices, r.missing_parent_edges, r.wrong_parent_distances, r.first_vertex) == \
            (inval, missing, wrong, first), trial
