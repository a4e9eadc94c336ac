"""Autotuner (cig_autotune): picks a launch shape from its candidate sets,
never slower than where it started, and the engine keeps bit-exact results
with the chosen shape through later batches."""
import pytest

import coincgraph_amd as cg
from tests.helpers import check_traversal, oracle_apply, oracle_graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if cg.device_count() < 1:
        pytest.skip("no HIP device")


@pytest.mark.parametrize("algo", ["sssp", "cc"])
def test_autotune_keeps_results(algo):
    print(f"[gpu-case] autotune {algo}", flush=True)
    g = cg.rmat_graph(12, 16, seed=71, symmetric=algo == "cc")
    s = cg.rmat_stream(g, 12, 3, 500, 500, seed=72, symmetric=algo == "cc")
    og = oracle_graph(g)
    with cg.Engine(algo, weight_mode=cg.WEIGHT_HASH64, weight_seed=2) as e:
        e.load_graph(g)
        r = e.autotune(0)
        assert r["edges_per_lane"] in (4, 8, 16) and r["split_degree"] in (32, 64, 128, 256)
        assert r["tile_len"] in (256, 512, 1024) and r["block_size"] in (256, 512)
        assert r["round_blocks_per_cu"] in (1, 2, 4) and r["trials"] > 10
        assert r["best_ms"] <= r["first_ms"]
        check_traversal(e, og, algo, 0, cg.WEIGHT_HASH64, 2, where="after autotune")
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            e.apply_batch(dels, adds)
            oracle_apply(og, dels, adds)
            check_traversal(e, og, algo, 0, cg.WEIGHT_HASH64, 2, where=f"tuned batch {b}")


@pytest.mark.parametrize("algo,adapt", [("sssp", 1), ("sssp", 4096), ("bfs", 16384), ("cc", 1), ("cc", 4096)])
def test_adaptive_split_keeps_results(algo, adapt):
    """cig_config.adapt_split only moves rows between in-round expansion and
    next-round tiles: every setting (1 = off) gives the oracle's values after
    the initial fixed point and after every batch, with long rows present
    (scale 14 hubs hold thousands of slots)."""
    print(f"[gpu-case] adapt_split {algo} {adapt}", flush=True)
    g = cg.rmat_graph(14, 16, seed=73, symmetric=algo == "cc")
    s = cg.rmat_stream(g, 14, 4, 1500, 1500, seed=74, symmetric=algo == "cc")
    og = oracle_graph(g)
    with cg.Engine(algo, weight_mode=cg.WEIGHT_HASH64, weight_seed=4, adapt_split=adapt) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        check_traversal(e, og, algo, 0, cg.WEIGHT_HASH64, 4, where=f"adapt {adapt} initial")
        e.load_updates(s)
        for b in range(s.nbatches):
            e.apply_next_batch()
            oracle_apply(og, *s.batch(b))
            check_traversal(e, og, algo, 0, cg.WEIGHT_HASH64, 4, where=f"adapt {adapt} batch {b}")
