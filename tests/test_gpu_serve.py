"""Serving loop (cig_serve_*): requests submitted together are coalesced
(update groups, merged multi-source queries) and every result equals the
result of the same request served alone at the same position of the
submission order. Admission by free HBM refuses what cannot fit."""
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


PLAN = ["U", ("sssp", [0, 5]), "U", "U", ("bfs", [3]), ("sssp", [7, 0, 9]), "U", "U", "U",
        ("sssp", [0]), "U", "U", ("bfs", [0, 1]), ("bfs", list(range(40)))]


@pytest.mark.parametrize("budget", [0.0, 0.3])
def test_serving_alone_equals_mixed(budget):
    """Updates keep their order; queries may be served after updates submitted
    later (they report the version they ran at). Each query's result equals
    the same query served alone at that version, and the final state equals
    applying every update in order. With a budget, host-tier rows are read
    zero-copy and the cache is refreshed between groups (refresh_every=1)."""
    print(f"[gpu-case] serve mixed budget={budget}", flush=True)
    g = cg.rmat_graph(11, 16, seed=61)
    nupd = sum(1 for p in PLAN if p == "U")
    s = cg.rmat_stream(g, 11, nupd, 400, 400, seed=62)
    kw = dict(weight_mode=cg.WEIGHT_HASH64, weight_seed=4, cache_bytes=int(g.m * 4 * 2 * budget))
    # mixed: everything submitted at once to one server
    with cg.Engine("sssp", **kw) as e:
        e.load_graph(g)
        if budget:
            assert e.info()["host_pool_used"] > 0
        e.init_graph(0)
        e.start()
        with cg.Server(e, max_group=4, refresh_every=1) as srv:
            tickets, b = [], 0
            for p in PLAN:
                if p == "U":
                    tickets.append(srv.submit_update(*s.batch(b)))
                    b += 1
                else:
                    tickets.append(srv.submit_query(p[0], p[1]))
            mixed = [srv.wait(t) for t in tickets]
        final = e.values().copy()
    assert all(r["status"] == 0 for r in mixed), [r["status"] for r in mixed]
    assert max(r["group_size"] for r in mixed) > 1, "nothing was coalesced"
    upd = [r["version"] for p, r in zip(PLAN, mixed) if p == "U"]
    assert upd == list(range(1, nupd + 1)), "updates must be applied in submission order"
    # alone: the same requests one at a time on a fresh engine, each query at its version
    og = oracle_graph(g)
    queries = sorted(((r["version"], i) for i, (p, r) in enumerate(zip(PLAN, mixed)) if p != "U"))
    with cg.Engine("sssp", **kw) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        applied = 0
        for ver, i in queries:
            while applied < ver:
                dels, adds = s.batch(applied)
                e.apply_batch(dels, adds)
                oracle_apply(og, dels, adds)
                applied += 1
            p = PLAN[i]
            alone = e.query(p[0], p[1])
            assert np.array_equal(mixed[i]["result"], alone), f"query {p} differs from alone at version {ver}"
        while applied < nupd:
            dels, adds = s.batch(applied)
            e.apply_batch(dels, adds)
            oracle_apply(og, dels, adds)
            applied += 1
        assert np.array_equal(final, e.values())
        check_traversal(e, og, "sssp", 0, cg.WEIGHT_HASH64, 4, where="after serving")


def test_serving_queries_do_not_hold_updates():
    """A query waiting behind nothing runs; updates submitted while queries
    wait are served first (their versions precede the queries'), unless a
    query has waited max_query_wait_ms."""
    print("[gpu-case] serve update priority", flush=True)
    g = cg.rmat_graph(10, 8, seed=65)
    s = cg.rmat_stream(g, 10, 6, 100, 100, seed=66)
    with cg.Engine("bfs") as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        with cg.Server(e, max_group=1, max_query_wait_ms=60_000) as srv:
            t = [srv.submit_query("bfs", [0])]
            t += [srv.submit_update(*s.batch(b)) for b in range(6)]
            t += [srv.submit_query("bfs", [1])]
            res = [srv.wait(x) for x in t]
        assert all(r["status"] == 0 for r in res)
        # the second query was submitted after every update: it sees all of them
        assert res[-1]["version"] == 6
        assert [r["version"] for r in res[1:7]] == list(range(1, 7))


def test_serving_admission_by_free_hbm():
    print("[gpu-case] serve admission", flush=True)
    g = cg.rmat_graph(9, 8, seed=63)
    s = cg.rmat_stream(g, 9, 1, 50, 50, seed=64)
    with cg.Engine("bfs") as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        before = e.values().copy()
        with cg.Server(e, hbm_reserve=1 << 62) as srv:
            u = srv.wait(srv.submit_update(*s.batch(0)))
            q = srv.wait(srv.submit_query("bfs", [0]))
        assert u["status"] == cg.ENOMEM and q["status"] == cg.ENOMEM
        assert np.array_equal(e.values(), before)
