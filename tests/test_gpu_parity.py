"""GPU parity: the HIP engine (through the C-ABI) against the CPU oracle.

BFS depths, SSSP distances and CC labels must be bit-exact after the initial
fixed point and after every update batch; parents must be valid witnesses;
the live topology must hold the same per-row multiset as the oracle's
reference-semantics chunk store. PageRank: normalised L1 (sum|x-x*|/sum|x*|)
against an FP64 Jacobi solve must stay <= 1e-6 with epsilon = 1e-7, and the
reference's own PageRankCheck invariant must hold.
"""
import os

import numpy as np
import pytest

import coincgraph_amd as cg
from oracle import oracle as O
from tests.helpers import (check_incoming, check_topology, check_traversal, oracle_apply,
                           oracle_graph, pr_errors)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if cg.device_count() < 1:
        pytest.skip("no HIP device")


@pytest.mark.parametrize("algo,wmode", [("sssp", cg.WEIGHT_REF128), ("sssp", cg.WEIGHT_HASH64),
                                        ("bfs", 0), ("cc", 0)])
@pytest.mark.parametrize("scale", [8, 14])
def test_converge_matches_oracle(algo, wmode, scale):
    _announce(f"converge {algo} w{wmode} s{scale}")
    g = cg.rmat_graph(scale, 16, seed=scale)
    og = oracle_graph(g)
    with cg.Engine(algo, weight_mode=wmode, weight_seed=99) as e:
        e.load_graph(g)
        e.init_graph(0)
        st = e.start()
        assert st["rounds"] > 0 and st["edges_scanned"] > 0
        check_traversal(e, og, algo, 0, wmode, 99, where="initial")


@pytest.mark.parametrize("algo,wmode", [("sssp", cg.WEIGHT_REF128), ("sssp", cg.WEIGHT_HASH64),
                                        ("bfs", 0), ("cc", 0)])
@pytest.mark.parametrize("staged,repair", [(False, "pull"), (True, "pull"), (False, "scan")])
def test_incremental_batches(algo, wmode, staged, repair):
    """Deletion repair by pulling over the device incoming-row index (default)
    or by scanning every row (CFG_SCAN_REPAIR) must give the same state."""
    _announce(f"batches {algo} w{wmode} staged={staged} repair={repair}")
    scale = 12
    sym = algo == "cc"
    g = cg.rmat_graph(scale, 16, seed=5, symmetric=sym)
    s = cg.rmat_stream(g, scale, 6, 400, 400, seed=6, symmetric=sym)
    og = oracle_graph(g)
    flags = cg.CFG_SCAN_REPAIR if repair == "scan" else 0
    with cg.Engine(algo, weight_mode=wmode, weight_seed=7, flags=flags) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            st = e.apply_batch(dels, adds, staged=staged)
            miss, inval = oracle_apply(og, dels, adds)
            assert st["missing_deletes"] == miss and st["invalid_updates"] == inval
            assert st["deletes_applied"] == len(dels) - miss
            check_traversal(e, og, algo, 0, wmode, 7, where=f"batch {b}")
        check_topology(e, og, "final")
        if repair == "pull":
            check_incoming(e, og, "final")


@pytest.mark.parametrize("algo,staged", [("cc", False), ("cc", True), ("sssp", False), ("bfs", False)])
def test_symmetric_mode(algo, staged):
    """CFG_SYMMETRIC on an undirected graph with symmetric batches: the
    outgoing rows serve as the incoming rows (no mirrored updates), and the
    state, topology and incoming rows equal the default engine's oracle."""
    _announce(f"symmetric {algo} staged={staged}")
    scale = 12
    g = cg.rmat_graph(scale, 16, seed=21, symmetric=True)
    s = cg.rmat_stream(g, scale, 6, 400, 400, seed=22, symmetric=True)
    og = oracle_graph(g)
    with cg.Engine(algo, weight_mode=cg.WEIGHT_HASH64, weight_seed=7, flags=cg.CFG_SYMMETRIC) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            st = e.apply_batch(dels, adds, staged=staged)
            miss, inval = oracle_apply(og, dels, adds)
            assert st["missing_deletes"] == miss and st["invalid_updates"] == inval
            assert st["deletes_applied"] == len(dels) - miss
            check_traversal(e, og, algo, 0, cg.WEIGHT_HASH64, 7, where=f"batch {b}")
        check_topology(e, og, "final")
        check_incoming(e, og, "final")


def test_symmetric_mode_pagerank():
    """PageRank on an undirected graph with CFG_SYMMETRIC gathers along the
    outgoing rows; ranks match the FP64 oracle."""
    _announce("symmetric pagerank")
    eps = 1e-7
    g = cg.rmat_graph(12, 16, seed=23, symmetric=True)
    s = cg.rmat_stream(g, 12, 3, 300, 300, seed=24, symmetric=True)
    og = oracle_graph(g)
    with cg.Engine("pr", pr_epsilon=eps, pr_max_rounds=2000, flags=cg.CFG_SYMMETRIC) as e:
        e.load_graph(g)
        e.init_graph()
        st = e.start()
        assert st["pull_rounds"] > 0
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            st = e.apply_batch(dels, adds)
            oracle_apply(og, dels, adds)
            assert st["converged"] == 1
            ref, _ = og.pagerank()
            l1n, maxerr = pr_errors(e.values(), ref)
            assert l1n <= 1e-6, f"batch {b}: normalised L1 {l1n:.3e} (max {maxerr:.3e})"


def test_symmetric_mode_rejects_directed_graph():
    g = cg.rmat_graph(10, 8, seed=25, symmetric=False)
    with cg.Engine("cc", flags=cg.CFG_SYMMETRIC) as e:
        with pytest.raises(cg.CigError):
            e.load_graph(g)


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc"])
def test_delete_then_insert_calls(algo):
    """del_edge and add_edge as separate calls (the reference's two stages,
    with the post-delete fixed point checked in between)."""
    _announce(f"del/add calls {algo}")
    g = cg.rmat_graph(11, 8, seed=21)
    s = cg.rmat_stream(g, 11, 4, 300, 300, seed=22)
    og = oracle_graph(g)
    with cg.Engine(algo) as e:
        e.load_graph(g)
        e.init_graph(3)
        e.start()
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            e.del_edge(dels)
            og.apply_deletes(dels["u"], dels["v"])
            check_traversal(e, og, algo, 3, where=f"delete stage {b}")
            e.add_edge(adds)
            og.apply_adds(adds["u"], adds["v"])
            check_traversal(e, og, algo, 3, where=f"batch {b}")


def _multigraph_case():
    """Rows of every delete class with parallel edges: a long out-row (vertex
    0, tiled scan), a medium out-row (vertex 3, wave per delete), a long and
    a medium incoming row (vertices 5 and 9), tiny rows (lane per delete);
    self-loops, many deletes on one row, missing and out-of-range records."""
    rng = np.random.default_rng(1234)
    n = 3000
    src = np.concatenate([np.zeros(6000, np.int64), np.full(2000, 3), rng.integers(0, n, 5000),
                          rng.integers(0, n, 1500), rng.integers(0, n, 20000)])
    dst = np.concatenate([rng.integers(1, n, 6000), rng.integers(0, 40, 2000), np.full(5000, 5),
                          np.full(1500, 9), rng.integers(0, n, 20000)])
    g_src = src.astype(np.uint32)
    g_dst = dst.astype(np.uint32)
    order = np.argsort(g_src, kind="stable")
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(g_src, minlength=n), out=off[1:])
    g = cg.HostGraph(n, off, g_dst[order])
    batches = []
    for b in range(4):
        pick = rng.integers(0, len(src), 2500)
        hub = rng.integers(0, 8000, 900)  # rows 0 and 3, repeated records included
        du = np.concatenate([src[pick], src[hub], rng.integers(0, n, 100), [n + 5, 1]])
        dv = np.concatenate([dst[pick], dst[hub], rng.integers(0, n, 100), [1, n + 9]])
        au = np.concatenate([rng.integers(0, n, 2000), np.zeros(3000, np.int64),
                             rng.integers(0, n, 700), [7, n + 1]])
        av = np.concatenate([rng.integers(0, n, 2000), rng.integers(0, n, 3000), np.full(700, 5),
                             [7, 2]])
        batches.append((cg.edges_array(du, dv), cg.edges_array(au, av)))
    return g, batches


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc"])
def test_multigraph_semantics(algo):
    _announce(f"multigraph {algo}")
    g, batches = _multigraph_case()
    og = oracle_graph(g)
    with cg.Engine(algo) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        for b, (dels, adds) in enumerate(batches):
            st = e.apply_batch(dels, adds)
            miss, inval = oracle_apply(og, dels, adds)
            assert (st["missing_deletes"], st["invalid_updates"]) == (miss, inval), b
            check_traversal(e, og, algo, 0, where=f"multigraph batch {b}")
            check_topology(e, og, f"multigraph batch {b}")
            check_incoming(e, og, f"multigraph batch {b}")


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc"])
def test_large_batch_medium_rows(algo):
    """A batch of more than 2^20 delete records (mirrors included) sends the
    medium rows (65-4096 slots) with 9-16 deletes to the 16-request claim
    pass; a smaller batch sends them to the long-row scan. Both give the
    oracle's results and topology."""
    _announce(f"large batch {algo}")
    g = cg.rmat_graph(18, 16, seed=91)
    s = cg.rmat_stream(g, 18, 2, 700000, 300000, seed=92)
    og = oracle_graph(g)
    # the case exists: medium rows with 9-16 deletes in the first batch
    deg = np.concatenate([np.diff(g.offsets).astype(np.int64), np.bincount(g.dst, minlength=g.n)])
    dels0, _ = s.batch(0)
    cnt = np.bincount(np.concatenate([dels0["u"].astype(np.int64), g.n + dels0["v"].astype(np.int64)]),
                      minlength=2 * g.n)
    assert np.count_nonzero((deg > 64) & (deg <= 4096) & (cnt > 8) & (cnt <= 16)) > 100
    with cg.Engine(algo) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            if b == 1:
                dels = dels[:100000]  # below the threshold: the plain form
            st = e.apply_batch(dels, adds)
            miss, inval = oracle_apply(og, dels, adds)
            assert (st["missing_deletes"], st["invalid_updates"]) == (miss, inval), b
            check_traversal(e, og, algo, 0, where=f"large batch {b}")
            check_topology(e, og, f"large batch {b}")


def test_empty_and_noop_batches():
    _announce("empty batches")
    g = cg.rmat_graph(10, 8, seed=3)
    og = oracle_graph(g)
    with cg.Engine("sssp") as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        empty = cg.edges_array([], [])
        e.apply_batch(empty, empty)
        e.apply_batch(cg.edges_array([5], [5]), empty)  # missing delete of a self-loop
        check_traversal(e, og, "sssp", 0, where="noop")


@pytest.mark.parametrize("budget_frac", [0.0, 0.3])
def test_zero_copy_tier_same_results(budget_frac):
    """Rows outside the HBM budget are read zero-copy from pinned host memory
    (and mutated there); results must not depend on placement."""
    _announce(f"zero-copy budget {budget_frac}")
    scale = 12
    g = cg.rmat_graph(scale, 16, seed=8)
    s = cg.rmat_stream(g, scale, 5, 500, 500, seed=9)
    og = oracle_graph(g)
    cap_bytes = max(1, int(budget_frac * g.m * 4 * 2)) if budget_frac else 4096
    with cg.Engine("sssp", cache_bytes=cap_bytes) as e:
        e.load_graph(g)
        info = e.info()
        assert info["host_pool_used"] > 0
        e.init_graph(0)
        st = e.start()
        assert st["host_edges_scanned"] > 0
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            e.apply_batch(dels, adds)
            oracle_apply(og, dels, adds)
            check_traversal(e, og, "sssp", 0, where=f"zc batch {b}")
            e.cache_refresh()
            check_traversal(e, og, "sssp", 0, where=f"zc refresh {b}")
            if budget_frac:
                # admission by activity: no host-tier row outscores an HBM row
                tier = e.gather(cg.FIELD_TIER).astype(bool)
                score = e.gather(cg.FIELD_HOTNESS)
                live = e.gather(cg.FIELD_DEGREE) > 0
                hbm, host = score[live & tier], score[live & ~tier]
                if hbm.size and host.size:
                    assert host.max() <= hbm.min(), (b, int(host.max()), int(hbm.min()))
        check_topology(e, og, "zc final")


def test_pool_overflow_recovery():
    """A budget just above the initial footprint forces relocation overflow,
    compaction and demotion to the host tier; results stay exact."""
    _announce("overflow recovery")
    scale = 11
    g = cg.rmat_graph(scale, 16, seed=10)
    og = oracle_graph(g)
    caps = sum(int(cg_cap) for cg_cap in
               [max(4, 1 << int(np.ceil(np.log2(max(d, 1))))) if d else 0
                for d in np.diff(g.offsets)])
    s = cg.rmat_stream(g, scale, 6, 100, 3000, seed=11)
    with cg.Engine("bfs", cache_bytes=int(caps * 4 * 1.15)) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            e.apply_batch(dels, adds)
            oracle_apply(og, dels, adds)
            check_traversal(e, og, "bfs", 0, where=f"overflow batch {b}")
        check_topology(e, og, "overflow final")


# PageRank round forms: "adaptive" (default: dense rounds gather along the
# incoming rows, sparse ones scatter), "push" (every round scatters),
# "pull" (every non-empty round gathers), "pull-host" (gathers with most
# incoming rows in the pinned host tier). Scale 14's vertex 0 has ~5.6k
# in-edges (several pull tiles); scale 10 has one-tile and short rows.
PR_FORMS = {
    "adaptive": dict(),
    "push": dict(flags=cg.CFG_PR_PUSH_ONLY),
    "pull": dict(pr_pull_divisor=0xFFFFFFFF),
    "pull-host": dict(pr_pull_divisor=0xFFFFFFFF, cache_bytes=None),
}


@pytest.mark.parametrize("form", list(PR_FORMS))
@pytest.mark.parametrize("scale", [10, 14])
def test_pagerank_static_and_batches(scale, form):
    _announce(f"pagerank s{scale} {form}")
    eps = 1e-7
    g = cg.rmat_graph(scale, 16, seed=12)
    s = cg.rmat_stream(g, scale, 4, 300, 300, seed=13)
    og = oracle_graph(g)
    kw = dict(PR_FORMS[form])
    if "cache_bytes" in kw:
        kw["cache_bytes"] = g.m * 4 * 2 // 3  # under a third of both row sets
    with cg.Engine("pr", pr_epsilon=eps, pr_max_rounds=2000, **kw) as e:
        e.load_graph(g)
        e.init_graph()
        st = e.start()
        if form.startswith("pull"):
            assert st["pull_rounds"] == st["rounds"] > 0
        if form == "pull-host":
            assert e.info()["host_pool_used"] > 0
        for b in range(-1, s.nbatches):
            if b >= 0:
                dels, adds = s.batch(b)
                st = e.apply_batch(dels, adds)
                oracle_apply(og, dels, adds)
            if form == "push":
                assert st["pull_rounds"] == 0
            assert st["converged"] == 1, f"batch {b}: not converged in 2000 rounds"
            ranks, resid = e.values(), e.residuals()
            ref, rounds = og.pagerank()
            l1n, maxerr = pr_errors(ranks, ref)
            assert l1n <= 1e-6, f"batch {b}: normalised L1 {l1n:.3e} (max {maxerr:.3e})"
            chk = og.pr_check(ranks, resid, eps)
            assert chk.finite and chk.max_residual <= np.float32(eps)
            assert chk.invariant_l1 <= 64 * np.finfo(np.float32).eps * max(1.0, chk.reference_mass)
            assert abs(st["pr_residual_l1"] - float(np.abs(resid.astype(np.float64)).sum())) <= \
                1e-6 * max(1.0, st["pr_residual_l1"])


def test_resident_stream_matches_calls():
    """get_update_file + per-batch apply (inputs resident in HBM) gives the
    same results as host-pointer batches."""
    _announce("resident stream")
    g = cg.rmat_graph(12, 16, seed=14)
    s = cg.rmat_stream(g, 12, 5, 500, 500, seed=15)
    og = oracle_graph(g)
    with cg.Engine("sssp", weight_mode=cg.WEIGHT_HASH64, weight_seed=5) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        e.load_updates(s)
        b = 0
        while e.batches_left():
            e.apply_next_batch()
            oracle_apply(og, *s.batch(b))
            check_traversal(e, og, "sssp", 0, cg.WEIGHT_HASH64, 5, where=f"stream {b}")
            b += 1
        assert b == s.nbatches


@pytest.mark.parametrize("algo", ["sssp", "cc", "pr"])
def test_cooperative_mode_matches(algo):
    """The persistent cooperative driver reaches the same fixed point as the
    chained round kernels."""
    _announce(f"cooperative {algo}")
    g = cg.rmat_graph(11, 16, seed=16, symmetric=algo == "cc")
    s = cg.rmat_stream(g, 11, 3, 200, 200, seed=17, symmetric=algo == "cc")
    og = oracle_graph(g)
    if algo == "pr":
        with cg.Engine("pr", flags=cg.CFG_COOPERATIVE, pr_epsilon=1e-7, pr_max_rounds=2000) as e:
            e.load_graph(g)
            e.init_graph()
            e.start()
            for b in range(s.nbatches):
                st = e.apply_batch(*s.batch(b))
                oracle_apply(og, *s.batch(b))
                assert st["converged"] == 1
                l1n, _ = pr_errors(e.values(), og.pagerank()[0])
                assert l1n <= 1e-6, (b, l1n)
        return
    with cg.Engine(algo, flags=cg.CFG_COOPERATIVE) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        for b in range(s.nbatches):
            e.apply_batch(*s.batch(b))
            oracle_apply(og, *s.batch(b))
            check_traversal(e, og, algo, 0, where=f"cooperative {b}")


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc", "pr"])
def test_golden_readme_example(algo):
    """README.md:27-35 update example on a hand-checked 5-vertex graph."""
    _announce(f"golden {algo}")
    g = cg.read_market(os.path.join(GOLDEN, "readme", "graph.txt"))
    s = cg.read_updates(os.path.join(GOLDEN, "readme", "updates.txt"),
                        os.path.join(GOLDEN, "readme", "batches.txt"))
    exp = np.loadtxt(os.path.join(GOLDEN, "readme", f"expected_{algo}.txt"), ndmin=2)
    with cg.Engine(algo, pr_epsilon=1e-7, pr_max_rounds=2000) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        e.apply_batch(*s.batch(0))
        got = e.values()
        if algo == "pr":
            np.testing.assert_allclose(got, exp[:, 1], rtol=2e-6, atol=1e-6)
        else:
            assert got.tolist() == exp[:, 1].astype(np.int64).tolist()


def test_sampled_round_timing():
    """kernel_timing_every = 3 times the round kernels of one call in three;
    the untimed calls report no launches, and every result stays exact."""
    _announce("sampled round timing")
    g = cg.rmat_graph(11, 16, seed=41)
    s = cg.rmat_stream(g, 11, 9, 200, 200, seed=42)
    og = oracle_graph(g)
    with cg.Engine("sssp", flags=cg.CFG_KERNEL_TIMING, kernel_timing_every=3) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        timed = 0
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            st = e.apply_batch(dels, adds)
            oracle_apply(og, dels, adds)
            check_traversal(e, og, "sssp", 0, where=f"timed batch {b}")
            assert st["total_ms"] > 0  # always timed
            if st["round_launches"]:
                timed += 1
                assert st["round_kernel_ms"] > 0 and st["mutate_ms"] > 0 and st["propagate_ms"] > 0
            else:  # the stage markers are skipped with the round timing
                assert st["round_kernel_ms"] == 0 and st["mutate_ms"] == 0
        assert timed == 3
