"""The five benchmark workloads at their own sizes, checked against the CPU
oracle after every batch (the reference's --check, run after each batch):

1. incremental BFS, R-MAT scale 16 edge factor 16, 10 x (1k deletes + 1k
   inserts): depths bit-exact and parents valid after every batch;
2. incremental SSSP, R-MAT scale 22, weights uniform in [1,64], 10 x (100k +
   100k), stream resident in HBM: CheckBellmanRelaxed + CheckParentWitness
   (hybrid_sssp.cu:211-316) after every batch, Dijkstra bit-exact at the end;
3. incremental min-label CC on a symmetrised R-MAT scale 24 (both directions
   stored, ~268M directed edges), 10 x (1M delete pairs + 1M insert pairs),
   undirected mode: CCCheck's flood (hybrid_cc_host.cu:8-32) after every batch;
4. incremental PageRank, R-MAT scale 25, alpha 0.85, epsilon 1e-7, 10 x
   (500k + 500k): PageRankCheck's invariant (hybrid_pr_host.cu:15-45) and
   max|r| <= epsilon after every batch, and the certificate
   ||x - x*||_1 <= (||r||_1 + defect) / 0.15 bounding the normalised L1
   error against the exact solution by 1e-6;
5. serving under an HBM edge budget (host-tier rows read zero-copy): update
   requests of log-uniform size interleaved with 64-source BFS/SSSP queries
   through cig_serve_*, every query checked against the oracle at the graph
   version it reports.
"""
import numpy as np
import pytest

import coincgraph_amd as cg
from oracle import oracle as O
from tests.helpers import check_traversal, oracle_apply, oracle_graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if cg.device_count() < 1:
        pytest.skip("no HIP device")


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.mark.timeout(600)
def test_config1_bfs_s16():
    _announce("config 1: BFS S16, 10 x 1k+1k")
    g = cg.rmat_graph(16, 16, seed=101)
    s = cg.rmat_stream(g, 16, 10, 1000, 1000, seed=102)
    og = oracle_graph(g)
    with cg.Engine("bfs") as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        check_traversal(e, og, "bfs", 0, where="initial")
        e.load_updates(s)
        for b in range(s.nbatches):
            st = e.apply_next_batch()
            miss, inval = oracle_apply(og, *s.batch(b))
            assert st["missing_deletes"] == miss == 0 and st["invalid_updates"] == inval == 0
            check_traversal(e, og, "bfs", 0, where=f"config 1 batch {b}")


@pytest.mark.timeout(900)
def test_config2_sssp_s22():
    _announce("config 2: SSSP S22, 10 x 100k+100k, weights [1,64]")
    seed_w = 2022
    g = cg.rmat_graph(22, 16, seed=201)
    s = cg.rmat_stream(g, 22, 10, 100_000, 100_000, seed=202)
    og = oracle_graph(g)
    with cg.Engine("sssp", weight_mode=cg.WEIGHT_HASH64, weight_seed=seed_w) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        e.load_updates(s)
        for b in range(s.nbatches):
            st = e.apply_next_batch()
            miss, _ = oracle_apply(og, *s.batch(b))
            assert st["missing_deletes"] == miss == 0
            assert st["deletes_applied"] == 100_000 and st["adds_applied"] == 100_000
            dist, par = e.values(), e.parents()
            bel = og.check_bellman(dist, 0, O.W_HASH64, seed_w)
            assert bel.relaxable_edges == 0 and bel.missing_tight_witnesses == 0, \
                f"batch {b}: relaxable {bel.relaxable_edges} (first {bel.first_src}->{bel.first_dst}), " \
                f"no tight witness {bel.missing_tight_witnesses} (first {bel.first_missing_witness})"
            pc = og.check_parent(dist, par, 0, O.W_HASH64, seed_w)
            assert pc.invalid_vertices == 0, f"batch {b}: parent of {pc.first_vertex} invalid"
            _announce(f"config 2 batch {b} ok ({st['total_ms']:.3f} ms)")
        exp = og.sssp(0, O.W_HASH64, seed_w)
        bad = np.nonzero(e.values() != exp)[0]
        assert bad.size == 0, f"final: {bad.size} distances differ from Dijkstra, first {bad[:5]}"
        assert O.distance_checksum(e.values()) == O.distance_checksum(exp)


@pytest.mark.timeout(900)
def test_config3_cc_s24_symmetric():
    _announce("config 3: CC symmetrised S24, 10 x 1M+1M pairs")
    g = cg.rmat_graph(24, 8, seed=301, symmetric=True)
    s = cg.rmat_stream(g, 24, 10, 1_000_000, 1_000_000, seed=302, symmetric=True)
    og = oracle_graph(g)
    with cg.Engine("cc", flags=cg.CFG_SYMMETRIC) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        assert og.cc_check(e.values()) == 0, "initial labels"
        e.load_updates(s)
        for b in range(s.nbatches):
            st = e.apply_next_batch()
            miss, _ = oracle_apply(og, *s.batch(b))
            assert st["missing_deletes"] == miss == 0
            errors = og.cc_check(e.values())
            assert errors == 0, f"batch {b}: {errors} labels differ from CCCheck's flood"
            _announce(f"config 3 batch {b} ok ({st['total_ms']:.3f} ms)")


@pytest.mark.timeout(900)
def test_config4_pagerank_s25():
    _announce("config 4: PR S25, 10 x 500k+500k, eps 1e-7")
    eps = 1e-7
    g = cg.rmat_graph(25, 16, seed=401)
    s = cg.rmat_stream(g, 25, 10, 500_000, 500_000, seed=402)
    og = oracle_graph(g)
    with cg.Engine("pr", pr_epsilon=eps, pr_max_rounds=1000) as e:
        e.load_graph(g)
        e.init_graph()
        st = e.start()
        e.load_updates(s)
        for b in range(-1, s.nbatches):
            if b >= 0:
                st = e.apply_next_batch()
                oracle_apply(og, *s.batch(b))
            assert st["converged"] == 1, f"batch {b}: not converged"
            x, r = e.values(), e.residuals()
            finite, defect, max_res, mass = og.pr_invariant(x, r)
            assert finite, f"batch {b}: non-finite rank or residual"
            assert max_res <= np.float32(eps), f"batch {b}: max|r| {max_res:.3e}"
            assert defect <= 64 * np.finfo(np.float32).eps * max(1.0, mass), \
                f"batch {b}: invariant defect {defect:.3e} (mass {mass:.3e})"
            # (I - 0.85 P^T) x = 0.15 + (defect terms) - r, and the inverse has
            # 1-norm <= 1/0.15 (dangling rows drop mass), so the L1 distance to
            # the exact solution x* is at most (sum|r| + defect) / 0.15
            l1_bound = (float(np.abs(r.astype(np.float64)).sum()) + defect) / 0.15
            assert l1_bound / (mass - l1_bound) <= 1e-6, f"batch {b}: L1 bound {l1_bound:.3e}"
            _announce(f"config 4 batch {b} ok (bound {l1_bound / mass:.2e}, {st['rounds']} rounds)")


@pytest.mark.timeout(600)
def test_config5_serving_under_budget():
    """Scaled-down serving workload: R-MAT scale 16 edge factor 32, an HBM edge
    budget of about 30% of the row capacity, update requests of log-uniform
    size (half deletes, half inserts) and 64-source queries in one stream of
    requests; every query equals the oracle at its reported version."""
    _announce("config 5: serving S16 EF32 under a 30% budget")
    seed_w = 55
    g = cg.rmat_graph(16, 32, seed=501)
    rng = np.random.default_rng(502)
    nupd = 24
    sizes = np.exp(rng.uniform(np.log(100), np.log(4000), nupd)).astype(np.int64)
    s = cg.rmat_stream(g, 16, nupd, int(sizes.max()), int(sizes.max()), seed=503)
    batches = []
    for b, k in enumerate(sizes):
        dels, adds = s.batch(b)
        batches.append((dels[: k // 2].copy(), adds[: k - k // 2].copy()))
    plan = []
    for b in range(nupd):
        plan.append(("U", b))
        if b % 6 == 2:
            plan.append(("Q", "sssp" if (b // 6) % 2 == 0 else "bfs", rng.integers(0, g.n, 64)))
    budget = int(g.m * 4 * 2 * 0.3)
    with cg.Engine("sssp", weight_mode=cg.WEIGHT_HASH64, weight_seed=seed_w, cache_bytes=budget) as e:
        e.load_graph(g)
        assert e.info()["host_pool_used"] > 0, "the budget must leave rows in the host tier"
        e.init_graph(0)
        e.start()
        with cg.Server(e, max_group=8) as srv:
            tickets = []
            for p in plan:
                if p[0] == "U":
                    tickets.append(srv.submit_update(*batches[p[1]]))
                else:
                    tickets.append(srv.submit_query(p[1], p[2]))
            res = [srv.wait(t) for t in tickets]
        assert all(r["status"] == 0 for r in res), [r["status"] for r in res]
        final = e.values().copy()
        info = e.info()
    assert info["host_pool_used"] > 0
    # replay on the oracle in version order; check every query at its version
    og = oracle_graph(g)
    applied = 0
    order = sorted(range(len(plan)), key=lambda i: (res[i]["version"], plan[i][0] == "U"))
    for i in order:
        p, r = plan[i], res[i]
        if p[0] == "U":
            continue
        while applied < r["version"]:
            oracle_apply(og, *batches[applied])
            applied += 1
        for j in range(0, 64, 4):  # every fourth source of the 64
            src = int(p[2][j])
            exp = og.bfs(src) if p[1] == "bfs" else og.sssp(src, O.W_HASH64, seed_w)
            bad = np.nonzero(r["result"][j] != exp)[0]
            assert bad.size == 0, f"query {i} ({p[1]}) source {src} at version {r['version']}: {bad.size} differ"
    versions = sorted(res[i]["version"] for i in range(len(plan)) if plan[i][0] == "U")
    assert versions == list(range(1, nupd + 1))
    while applied < nupd:
        oracle_apply(og, *batches[applied])
        applied += 1
    exp = og.sssp(0, O.W_HASH64, seed_w)
    assert np.array_equal(final, exp), "incremental source-0 state after serving"
