"""PageRank engines number vertices internally by decreasing out-degree
(default) or keep the caller's ids (CIG_CFG_PR_KEEP_IDS). Every id that
crosses the ABI is the caller's either way: ranks, residuals, degrees, tiers,
topology and incoming rows equal the oracle's (or the keep-ids engine's), and
records, resident streams and served updates are taken in the caller's ids."""
import numpy as np
import pytest

import coincgraph_amd as cg
from oracle import oracle as O
from tests.helpers import check_incoming, check_topology, oracle_apply, oracle_graph, pr_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if cg.device_count() < 1:
        pytest.skip("no HIP device")


def _check_pr(e, og, eps, where):
    ranks, resid = e.values(), e.residuals()
    ref, _ = og.pagerank()
    l1n, maxerr = pr_errors(ranks, ref)
    assert l1n <= 1e-6, f"{where}: normalised L1 {l1n:.3e} (max {maxerr:.3e})"
    chk = og.pr_check(ranks, resid, eps)
    assert chk.finite and chk.max_residual <= np.float32(eps), where
    assert chk.invariant_l1 <= 64 * np.finfo(np.float32).eps * max(1.0, chk.reference_mass), where


@pytest.mark.parametrize("budget", [False, True])
def test_renumbered_pr_matches_keep_ids_and_oracle(budget):
    print(f"[gpu-case] pr renumber budget={budget}", flush=True)
    eps = 1e-7
    g = cg.rmat_graph(13, 16, seed=71)
    s = cg.rmat_stream(g, 13, 4, 400, 400, seed=72)
    og = oracle_graph(g)
    kw = dict(pr_epsilon=eps, pr_max_rounds=2000)
    if budget:
        kw["cache_bytes"] = g.m * 4 * 2 // 3
    with cg.Engine("pr", **kw) as r, cg.Engine("pr", flags=cg.CFG_PR_KEEP_IDS, **kw) as k:
        for e in (r, k):
            e.load_graph(g)
            e.init_graph()
            assert e.start()["converged"] == 1
        deg = np.diff(g.offsets.astype(np.int64))
        assert np.array_equal(r.gather(cg.FIELD_DEGREE), deg)
        if budget:
            assert r.info()["host_pool_used"] > 0
            # rows are placed by capacity, so a vertex's tier follows its degree class
            tr, tk = r.gather(cg.FIELD_TIER), k.gather(cg.FIELD_TIER)
            assert tr.sum() > 0 and (tr == 0).sum() > 0
            assert np.all(deg[tr == 1] > 0)
        _check_pr(r, og, eps, "renumbered initial")
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            for e in (r, k):
                assert e.apply_batch(dels, adds)["converged"] == 1
            oracle_apply(og, dels, adds)
            _check_pr(r, og, eps, f"renumbered batch {b}")
            _check_pr(k, og, eps, f"keep-ids batch {b}")
        check_topology(r, og, "renumbered final")
        check_incoming(r, og, "renumbered final")
        assert np.array_equal(r.gather(cg.FIELD_DEGREE), k.gather(cg.FIELD_DEGREE))
        with pytest.raises(cg.CigError):
            r.query("bfs", [0])
        assert k.query("bfs", [0]).shape[-1] == g.n


def test_renumbered_pr_resident_stream_and_groups():
    """A resident stream is taken in the caller's ids, one batch at a time and
    as coalesced groups."""
    print("[gpu-case] pr renumber stream", flush=True)
    eps = 1e-7
    g = cg.rmat_graph(12, 16, seed=73)
    s = cg.rmat_stream(g, 12, 6, 300, 300, seed=74)
    og = oracle_graph(g)
    with cg.Engine("pr", pr_epsilon=eps, pr_max_rounds=2000) as e:
        e.load_graph(g)
        e.init_graph()
        e.start()
        e.load_updates(s)
        e.apply_next_batch()
        oracle_apply(og, *s.batch(0))
        _check_pr(e, og, eps, "stream batch 0")
        e.apply_next_batches(5)
        for b in range(1, 6):
            oracle_apply(og, *s.batch(b))
        _check_pr(e, og, eps, "stream group 1-5")
        check_topology(e, og, "stream final")
