"""Incremental maintenance of up to 64 BFS / SSSP sources in one engine
(cig_init with several sources): after the initial fixed point and after
every update batch, each source's distances must equal the oracle's
single-source run bit for bit and its parents must be valid witnesses
(CheckParentWitness, hybrid_sssp.cu:266-316), on both tiers, in the fused,
staged, separate-call and batched forms, and equal to single-source engines
run side by side."""
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


def _check_all(e, og, algo, sources, wmode, seed, where, parents=True):
    om = O.W_UNIT if algo == "bfs" else (O.W_HASH64 if wmode == cg.WEIGHT_HASH64 else O.W_REF128)
    vals = e.values()
    par = e.parents() if parents else None
    assert vals.shape == (len(sources), e.n)
    for i, src in enumerate(sources):
        exp = og.bfs(int(src)) if algo == "bfs" else og.sssp(int(src), om, seed)
        bad = np.nonzero(vals[i] != exp)[0]
        assert bad.size == 0, (f"{where}: source {src} (#{i}): {bad.size} mismatches, first {bad[:5].tolist()} "
                               f"got {vals[i][bad[:5]].tolist()} exp {exp[bad[:5]].tolist()}")
        if parents:
            res = og.check_parent(vals[i], par[i], int(src), om, seed)
            assert res.invalid_vertices == 0, f"{where}: source {src}: parent of {res.first_vertex} invalid"


@pytest.mark.parametrize("algo,wmode", [("sssp", cg.WEIGHT_HASH64), ("sssp", cg.WEIGHT_REF128), ("bfs", 0)])
@pytest.mark.parametrize("k", [2, 7, 64])
def test_multisource_batches_match_oracle(algo, wmode, k):
    print(f"[gpu-case] multi-source {algo} w{wmode} k={k}", flush=True)
    g = cg.rmat_graph(12, 16, seed=81)
    s = cg.rmat_stream(g, 12, 5, 400, 400, seed=82)
    og = oracle_graph(g)
    rng = np.random.default_rng(k)
    sources = rng.integers(0, g.n, k).astype(np.uint32)
    sources[0] = 0
    if k > 2:
        sources[1] = sources[2]  # a repeated source is allowed
    with cg.Engine(algo, weight_mode=wmode, weight_seed=11) as e:
        e.load_graph(g)
        e.init_graph(sources)
        e.start()
        _check_all(e, og, algo, sources, wmode, 11, "initial")
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            st = e.apply_batch(dels, adds)
            miss, _ = oracle_apply(og, dels, adds)
            assert st["missing_deletes"] == miss
            _check_all(e, og, algo, sources, wmode, 11, f"batch {b}")


@pytest.mark.parametrize("form", ["staged", "separate", "grouped", "host-tier", "planted",
                                  "one-stream", "claims-on-engine-stream"])
def test_multisource_forms(form):
    """Staged batches, separate delete/insert calls, the batched (grouped) form,
    rows in the pinned host tier, a stream with duplicate, missing and
    out-of-range records, and batches whose grouping and claims run on the
    engine stream (n_stream 1) or whose claims do (n_stream 2) all keep every
    source exact."""
    print(f"[gpu-case] multi-source form {form}", flush=True)
    g = cg.rmat_graph(11, 16, seed=83)
    s = cg.rmat_stream(g, 11, 6, 300, 300, seed=84)
    og = oracle_graph(g)
    sources = np.array([0, 1, 17, 400, 1023, 2047, 5, 5, 9, 33], dtype=np.uint32)
    kw = dict(weight_mode=cg.WEIGHT_HASH64, weight_seed=12)
    if form == "host-tier":
        kw["cache_bytes"] = g.m * 4 * 2 // 3
    if form == "one-stream":
        kw["n_stream"] = 1
    if form == "claims-on-engine-stream":
        kw["n_stream"] = 2
    with cg.Engine("sssp", **kw) as e:
        e.load_graph(g)
        if form == "host-tier":
            assert e.info()["host_pool_used"] > 0
        e.init_graph(sources)
        e.start()
        if form == "grouped":
            e.load_updates(s)
            e.apply_next_batches(3)
            for b in range(3):
                oracle_apply(og, *s.batch(b))
            _check_all(e, og, "sssp", sources, cg.WEIGHT_HASH64, 12, "group 0-2")
            e.apply_next_batches(3)
            for b in range(3, 6):
                oracle_apply(og, *s.batch(b))
            _check_all(e, og, "sssp", sources, cg.WEIGHT_HASH64, 12, "group 3-5")
            return
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            if form == "planted":
                # duplicates of existing edges, a delete of a missing edge, out-of-range ids
                dels = np.concatenate([dels, dels[:3], np.array([(1, 2, 0), (g.n + 5, 0, 0)],
                                                                dtype=cg.EDGE_DTYPE)])
                adds = np.concatenate([adds, adds[:5], np.array([(0, g.n, 0)], dtype=cg.EDGE_DTYPE)])
            if form == "staged":
                e.apply_batch(dels, adds, staged=True)
            elif form == "separate":
                e.del_edge(dels)
                e.add_edge(adds)
            else:
                e.apply_batch(dels, adds)
            oracle_apply(og, dels, adds)
            _check_all(e, og, "sssp", sources, cg.WEIGHT_HASH64, 12, f"{form} batch {b}")


def test_multisource_equals_single_source_engines():
    """Source by source, the multi-source engine equals a single-source engine
    fed the same batches (values and parents' validity), and initialising it
    again with one source returns to the single-source state."""
    print("[gpu-case] multi-source vs single engines", flush=True)
    g = cg.rmat_graph(11, 16, seed=85)
    s = cg.rmat_stream(g, 11, 4, 300, 300, seed=86)
    sources = [0, 3, 77]
    kw = dict(weight_mode=cg.WEIGHT_HASH64, weight_seed=13)
    singles = [cg.Engine("sssp", **kw) for _ in sources]
    try:
        with cg.Engine("sssp", **kw) as e:
            e.load_graph(g)
            e.init_graph(sources)
            e.start()
            for one, src in zip(singles, sources):
                one.load_graph(g)
                one.init_graph(src)
                one.start()
            for b in range(s.nbatches):
                e.apply_batch(*s.batch(b))
                vals = e.values()
                for i, one in enumerate(singles):
                    one.apply_batch(*s.batch(b))
                    assert np.array_equal(vals[i], one.values()), (b, sources[i])
            e.init_graph(77)
            e.start()
            assert e.values().shape == (g.n,)
            assert np.array_equal(e.values(), singles[2].values())
    finally:
        for one in singles:
            one.close()


def test_multisource_rejects_bad_arguments():
    print("[gpu-case] multi-source arguments", flush=True)
    g = cg.rmat_graph(8, 8, seed=87)
    with cg.Engine("sssp") as e:
        e.load_graph(g)
        with pytest.raises(cg.CigError) as ei:
            e.init_graph(list(range(65)))
        assert ei.value.status == cg.EINVAL
        with pytest.raises(cg.CigError):
            e.init_graph([0, g.n])
    with cg.Engine("sssp", flags=cg.CFG_SCAN_REPAIR) as e:
        e.load_graph(g)
        with pytest.raises(cg.CigError) as ei:
            e.init_graph([0, 1])
        assert ei.value.status == cg.EINVAL
