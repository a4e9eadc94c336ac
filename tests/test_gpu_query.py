"""Multi-source queries (cig_query): up to 64 BFS / SSSP sources at once on
the live graph must equal the oracle's single-source runs bit for bit, on
both tiers, after update batches, and must leave the engine's incremental
state unchanged."""
import numpy as np
import pytest

import coincgraph_amd as cg
from oracle import oracle as O
from tests.helpers import oracle_apply, oracle_graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if cg.device_count() < 1:
        pytest.skip("no HIP device")


def _expect(og, algo, src, wmode, seed):
    if algo == "bfs":
        return og.bfs(int(src))
    return og.sssp(int(src), O.W_HASH64 if wmode == cg.WEIGHT_HASH64 else O.W_REF128, seed)


@pytest.mark.parametrize("algo", ["bfs", "sssp"])
@pytest.mark.parametrize("nsrc", [1, 7, 64])
@pytest.mark.parametrize("budget", [0.0, 0.3])
def test_query_matches_oracle(algo, nsrc, budget):
    print(f"[gpu-case] query {algo} k={nsrc} budget={budget}", flush=True)
    g = cg.rmat_graph(11, 16, seed=51)
    s = cg.rmat_stream(g, 11, 2, 300, 300, seed=52)
    og = oracle_graph(g)
    rng = np.random.default_rng(nsrc)
    cache = int(g.m * 4 * budget) if budget else 0
    with cg.Engine("sssp", weight_mode=cg.WEIGHT_HASH64, weight_seed=9, cache_bytes=cache) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        for b in range(s.nbatches + 1):
            before = e.values().copy()
            sources = rng.integers(0, g.n, nsrc)
            sources[0] = 0
            got = e.query(algo, sources)
            assert got.shape == (nsrc, g.n)
            for i, src in enumerate(sources):
                exp = _expect(og, algo, src, cg.WEIGHT_HASH64, 9)
                bad = np.nonzero(got[i] != exp)[0]
                assert bad.size == 0, f"batch {b} source {src}: {bad.size} mismatches at {bad[:5]}"
            assert np.array_equal(e.values(), before), "query changed the engine state"
            if algo == "sssp":  # source 0 alone in the mix equals the incremental engine
                assert np.array_equal(got[0], before)
            if b < s.nbatches:
                dels, adds = s.batch(b)
                e.apply_batch(dels, adds)
                oracle_apply(og, dels, adds)


def test_query_rejects_bad_arguments():
    g = cg.rmat_graph(8, 4, seed=1)
    with cg.Engine("bfs") as e:
        e.load_graph(g)
        with pytest.raises(cg.CigError):
            e.query("bfs", np.arange(65))
        with pytest.raises(cg.CigError):
            e.query("cc", [0])
        with pytest.raises(cg.CigError):
            e.query("bfs", [g.n])


@pytest.mark.parametrize("algo", ["bfs", "sssp"])
def test_large_result_copy_matches(algo):
    """A result above 256 MiB leaves the device through pinned bounce buffers
    in chunks; it must equal the direct copy of smaller queries from the same
    sources (64 sources x 2^21 vertices x 4 B = two 256-MiB chunks)."""
    print(f"[gpu-case] large query result {algo}", flush=True)
    g = cg.rmat_graph(21, 2, seed=53)
    with cg.Engine("sssp", weight_mode=cg.WEIGHT_HASH64, weight_seed=9) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        sources = (np.arange(64, dtype=np.uint64) * 16411 % g.n).astype(np.uint32)
        sources[0] = 0
        big = e.query(algo, sources)
        assert big.nbytes > 256 << 20
        for lo in (0, 32, 60):
            small = e.query(algo, sources[lo:lo + 4])
            assert np.array_equal(big[lo:lo + 4], small), f"sources {lo}..{lo + 3}"
        if algo == "sssp":
            assert np.array_equal(big[0], e.values())
        ref = _expect(oracle_graph(g), algo, sources[1], cg.WEIGHT_HASH64, 9)
        assert np.array_equal(big[1], ref)
