"""Graph-replayed batches (CIG_CFG_GRAPH_BATCHES): the fixed-length parts of
every batch are replayed from two captured HIP graphs; results must equal
the oracle after every replay, with the program's own eager work (gathers,
other engines' batches, synchronisations) between replays, and a pool
compaction (new buffers) must lead to a fresh capture, not a stale replay."""
import numpy as np
import pytest

import coincgraph_amd as cg
from tests.helpers import check_incoming, check_topology, check_traversal, oracle_apply, oracle_graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if cg.device_count() < 1:
        pytest.skip("no HIP device")


@pytest.mark.parametrize("algo,wmode", [("sssp", cg.WEIGHT_HASH64), ("bfs", 0), ("cc", 0)])
@pytest.mark.parametrize("resident", [True, False])
@pytest.mark.parametrize("n_stream", [3, 2, 1])  # 2, 3: grouping graph on the second stream
def test_graph_batches_match_oracle(algo, wmode, resident, n_stream):
    print(f"[gpu-case] graph batches {algo} resident={resident} n_stream={n_stream}", flush=True)
    sym = algo == "cc"
    g = cg.rmat_graph(12, 16, seed=111, symmetric=sym)
    s = cg.rmat_stream(g, 12, 8, 400, 400, seed=112, symmetric=sym)
    og = oracle_graph(g)
    flags = cg.CFG_GRAPH_BATCHES | (cg.CFG_SYMMETRIC if sym else 0)
    with cg.Engine(algo, weight_mode=wmode, weight_seed=31, flags=flags, n_stream=n_stream) as e, \
            cg.Engine(algo, weight_mode=wmode, weight_seed=31, flags=flags & ~cg.CFG_GRAPH_BATCHES) as eager:
        for x in (e, eager):
            x.load_graph(g)
            x.init_graph(0)
            x.start()
        if resident:
            e.load_updates(s)
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            st = e.apply_next_batch() if resident else e.apply_batch(dels, adds)
            eager.apply_batch(dels, adds)  # eager work on the same device between replays
            miss, inval = oracle_apply(og, dels, adds)
            assert st["missing_deletes"] == miss and st["invalid_updates"] == inval
            check_traversal(e, og, algo, 0, wmode, 31, where=f"graph batch {b}")
            assert np.array_equal(e.values(), eager.values()), b
        check_topology(e, og, "graph final")


def test_graph_batches_recapture_after_compaction():
    """A pool too small for the inserts forces compactions (new pool buffers):
    the next batches must capture new graphs and stay exact."""
    print("[gpu-case] graph recapture", flush=True)
    g = cg.rmat_graph(11, 16, seed=113)
    s = cg.rmat_stream(g, 11, 6, 100, 3000, seed=114)  # many more inserts than deletes
    og = oracle_graph(g)
    caps = sum(max(4, 1 << int(np.ceil(np.log2(max(d, 1))))) if d else 0 for d in np.diff(g.offsets))
    with cg.Engine("sssp", flags=cg.CFG_GRAPH_BATCHES, cache_bytes=int(caps * 4 * 1.15)) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        comp = 0
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            st = e.apply_batch(dels, adds)
            comp = max(comp, st["pool_compactions"])
            oracle_apply(og, dels, adds)
            check_traversal(e, og, "sssp", 0, where=f"recapture batch {b}")
        assert comp > 0, "the stream never overflowed the pool"
        check_topology(e, og, "recapture final")


@pytest.mark.parametrize("tier", ["hbm", "host"])
def test_in_place_compaction_keeps_results(tier):
    """Relocation garbage is packed away in place (HBM: device kernels; host
    tier: host threads), by cig_compact between batches and by a tail that
    overflows inside a batch (a pool with almost no slack, as many deletes
    as inserts); every result stays exact."""
    print(f"[gpu-case] in-place compaction {tier}", flush=True)
    g = cg.rmat_graph(11, 16, seed=115)
    s = cg.rmat_stream(g, 11, 16, 1500, 1500, seed=116)
    og = oracle_graph(g)
    kw = dict(weight_mode=cg.WEIGHT_HASH64, weight_seed=3, pool_slack=64)
    if tier == "host":
        kw["cache_bytes"] = g.m * 4 // 5  # most rows of both directions in the host tier
    with cg.Engine("sssp", **kw) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            e.apply_batch(dels, adds)
            oracle_apply(og, dels, adds)
            check_traversal(e, og, "sssp", 0, cg.WEIGHT_HASH64, 3, where=f"{tier} batch {b}")
            if b in (5, 11):
                before = e.info()
                e.compact(tier)
                after = e.info()
                assert after["inplace_compactions"] == before["inplace_compactions"] + 1, (before, after)
                used = "host_pool_used" if tier == "host" else "hbm_pool_used"
                assert after[used] <= before[used], (before, after)
                check_traversal(e, og, "sssp", 0, cg.WEIGHT_HASH64, 3, where=f"{tier} compacted after {b}")
                check_topology(e, og, f"{tier} compacted after {b}")
        info = e.info()
        if tier == "host":
            assert info["host_compactions"] >= 2, info
        check_topology(e, og, f"{tier} final")


def test_host_tier_free_blocks_reused():
    """Relocated host-tier rows and rows loaded into HBM free their host
    blocks; later relocations and evictions of the same size class take them
    (free-block lists) instead of growing the tail. Results and topology stay
    exact through batches, refreshes that move rows both ways, and packing."""
    print("[gpu-case] host free blocks", flush=True)
    g = cg.rmat_graph(12, 16, seed=117)
    s = cg.rmat_stream(g, 12, 18, 800, 1200, seed=118)
    og = oracle_graph(g)
    with cg.Engine("sssp", weight_mode=cg.WEIGHT_HASH64, weight_seed=5, cache_bytes=g.m * 4 // 2,
                   pool_slack=64) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        moved = 0
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            e.apply_batch(dels, adds)
            oracle_apply(og, dels, adds)
            check_traversal(e, og, "sssp", 0, cg.WEIGHT_HASH64, 5, where=f"free-list batch {b}")
            if b % 3 == 2:
                moved += e.cache_refresh()["relocated_rows"]
                check_traversal(e, og, "sssp", 0, cg.WEIGHT_HASH64, 5, where=f"free-list refresh {b}")
            if b == 10:
                e.compact("host")
        assert moved > 0, "no refresh moved a row"
        check_topology(e, og, "free-list final")
        check_incoming(e, og, "free-list final")
