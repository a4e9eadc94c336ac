"""Batched form: consecutive update batches applied as one group
(cig_apply_next_batches) must leave exactly the state of applying them one
by one: bit-exact values, valid parents, the same topology and incoming-row
multisets, and the same applied / missing / invalid counts. The stream is
built so that later batches delete edges that earlier batches of the group
added (and re-add deleted ones), with parallel edges and missing records."""
import numpy as np
import pytest

import coincgraph_amd as cg
from tests.helpers import (check_incoming, check_topology, check_traversal, oracle_apply,
                           oracle_graph, pr_errors)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if cg.device_count() < 1:
        pytest.skip("no HIP device")


def _conflict_stream(g, nb, size, seed):
    rng = np.random.default_rng(seed)
    src = np.repeat(np.arange(g.n), np.diff(g.offsets).astype(np.int64))
    adds, dels, sizes = [], [], []
    prev_adds = None
    prev_dels = None
    for b in range(nb):
        pick = rng.integers(0, len(src), size)
        du, dv = src[pick], g.dst[pick]
        if prev_adds is not None:  # delete some of the previous batch's adds (twice for a few)
            k = size // 3
            du = np.concatenate([du, prev_adds[0][:k], prev_adds[0][:8]])
            dv = np.concatenate([dv, prev_adds[1][:k], prev_adds[1][:8]])
        du = np.concatenate([du, rng.integers(0, g.n, 20), [g.n + 3]])
        dv = np.concatenate([dv, rng.integers(0, g.n, 20), [0]])
        au = rng.integers(0, g.n, size)
        av = rng.integers(0, g.n, size)
        if prev_dels is not None:  # re-add some edges the previous batch deleted
            k = size // 4
            au = np.concatenate([au, prev_dels[0][:k]])
            av = np.concatenate([av, prev_dels[1][:k]])
        au = np.concatenate([au, au[:10], [1]])  # parallel adds, one invalid
        av = np.concatenate([av, av[:10], [g.n + 7]])
        adds.append(cg.edges_array(au, av))
        dels.append(cg.edges_array(du, dv))
        sizes.append((len(au), len(du)))
        prev_adds, prev_dels = (au, av), (du, dv)
    return cg.UpdateStream(np.concatenate(adds), np.concatenate(dels),
                           np.asarray(sizes, dtype=np.uint32).reshape(-1, 2))


@pytest.mark.parametrize("algo,wmode", [("sssp", cg.WEIGHT_HASH64), ("bfs", 0), ("cc", 0)])
def test_grouped_batches_match_sequential(algo, wmode):
    print(f"[gpu-case] grouped {algo}", flush=True)
    g = cg.rmat_graph(12, 16, seed=41, symmetric=algo == "cc")
    s = _conflict_stream(g, 9, 600, 42)
    og = oracle_graph(g)
    with cg.Engine(algo, weight_mode=wmode, weight_seed=3) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        e.load_updates(s)
        b = 0
        for k in (1, 3, 5):
            st = e.apply_next_batches(k)
            miss = inval = applied = added = 0
            for j in range(b, b + k):
                dels, adds = s.batch(j)
                m, i = oracle_apply(og, dels, adds)
                miss += m
                inval += i
                applied += len(dels) - m - int(((dels["u"] >= g.n) | (dels["v"] >= g.n)).sum())
                added += len(adds) - int(((adds["u"] >= g.n) | (adds["v"] >= g.n)).sum())
            b += k
            assert (st["missing_deletes"], st["invalid_updates"]) == (miss, inval), k
            assert (st["deletes_applied"], st["adds_applied"]) == (applied, added), k
            check_traversal(e, og, algo, 0, wmode, 3, where=f"group of {k}")
            check_topology(e, og, f"group of {k}")
            check_incoming(e, og, f"group of {k}")
        assert e.batches_left() == 0


def test_grouped_pagerank_matches_fp64():
    print("[gpu-case] grouped pr", flush=True)
    g = cg.rmat_graph(11, 8, seed=43)
    s = _conflict_stream(g, 6, 300, 44)
    og = oracle_graph(g)
    with cg.Engine("pr", pr_epsilon=1e-7, pr_max_rounds=2000) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        e.load_updates(s)
        e.apply_next_batches(2)
        e.apply_next_batches(4)
        for j in range(6):
            oracle_apply(og, *s.batch(j))
        check_topology(e, og, "pr group")
        l1n, _ = pr_errors(e.values(), og.pagerank()[0])
        assert l1n <= 1e-6


@pytest.mark.parametrize("algo,wmode", [("sssp", cg.WEIGHT_HASH64), ("cc", 0)])
def test_each_form_matches_one_call_per_batch(algo, wmode):
    """cig_apply_next_batches with CIG_BATCHES_EACH is the host loop over the
    batches: per-batch statistics equal one call per batch, and the state
    after the call is the oracle's."""
    print(f"[gpu-case] each-form {algo}", flush=True)
    g = cg.rmat_graph(12, 16, seed=45, symmetric=algo == "cc")
    s = _conflict_stream(g, 6, 500, 46)
    og = oracle_graph(g)
    keys = ("deletes_applied", "adds_applied", "missing_deletes", "invalid_updates")
    with cg.Engine(algo, weight_mode=wmode, weight_seed=5) as a, \
            cg.Engine(algo, weight_mode=wmode, weight_seed=5) as b:
        for e in (a, b):
            e.load_graph(g)
            e.init_graph(0)
            e.start()
            e.load_updates(s)
        one = [b.apply_next_batch() for _ in range(6)]
        each = a.apply_next_batches(6, each=True)
        assert len(each) == 6
        for j in range(6):
            oracle_apply(og, *s.batch(j))
            assert [each[j][k] for k in keys] == [one[j][k] for k in keys], j
            assert each[j]["total_ms"] > 0
        assert np.array_equal(a.values(), b.values())
        check_traversal(a, og, algo, 0, wmode, 5, where="each form")
        check_topology(a, og, "each form")
        assert a.batches_left() == 0


@pytest.mark.parametrize("budget,n_stream", [(0.0, 3), (0.3, 3), (0.0, 2), (0.0, 1)])
def test_two_stream_batches_interleaved_with_other_calls(budget, n_stream):
    """A traversal batch groups its records on a second stream beside the
    invalidation waves and clears its counts there at the end; the next
    resident batch starts without waiting for the engine stream. Resident
    batches interleaved with host-pointer batches, staged batches, grouped
    batches, gathers, cache refreshes and a compaction must stay equal to the
    oracle after every step (n_stream = 1: one compute stream, the same
    results)."""
    print(f"[gpu-case] two-stream interleave budget={budget} n_stream={n_stream}", flush=True)
    g = cg.rmat_graph(12, 16, seed=47)
    s = _conflict_stream(g, 12, 400, 48)
    extra = _conflict_stream(g, 3, 300, 49)
    og = oracle_graph(g)
    cache = int(g.m * 4 * 2 * budget)
    with cg.Engine("sssp", weight_mode=cg.WEIGHT_HASH64, weight_seed=7, cache_bytes=cache,
                   n_stream=n_stream) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        e.load_updates(s)
        j = 0

        def resident(k=1, each=False):
            nonlocal j
            if each:
                e.apply_next_batches(k, each=True)
            elif k > 1:
                e.apply_next_batches(k)
            else:
                e.apply_next_batch()
            for b in range(j, j + k):
                oracle_apply(og, *s.batch(b))
            j += k

        resident()
        resident()  # consecutive resident batches (no fork)
        check_traversal(e, og, "sssp", 0, cg.WEIGHT_HASH64, 7, where="two resident")
        d, a = extra.batch(0)
        e.apply_batch(d, a)  # host records: copied on the engine stream first
        oracle_apply(og, d, a)
        resident()
        check_topology(e, og, "after host batch")
        e.cache_refresh()
        resident(2, each=True)
        d, a = extra.batch(1)
        e.apply_batch(d, a, staged=True)
        oracle_apply(og, d, a)
        resident()
        resident(3)  # grouped
        check_traversal(e, og, "sssp", 0, cg.WEIGHT_HASH64, 7, where="after grouped")
        e.compact(0)
        resident()
        d, a = extra.batch(2)
        e.apply_batch(d, a)
        oracle_apply(og, d, a)
        resident(2, each=True)
        check_traversal(e, og, "sssp", 0, cg.WEIGHT_HASH64, 7, where="end")
        check_topology(e, og, "end")
        check_incoming(e, og, "end")
        assert e.batches_left() == 0
