"""Alternative repair executors, each bit-exact against the oracle after every
batch: delta-ordered SSSP rounds (CIG_CFG_ORDERED_REPAIR, the reference's
ordered_gpu_repair.cuh:117-160) and union-find CC on undirected graphs
(CIG_CFG_CC_UNION_REPAIR, cc_union_repair.cuh:9-137)."""
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


@pytest.mark.parametrize("wmode", [cg.WEIGHT_HASH64, cg.WEIGHT_REF128])
@pytest.mark.parametrize("form,delta", [("fused", 0), ("staged", 0), ("grouped", 0), ("fused", 7), ("fused", 1)])
def test_ordered_repair_matches_oracle(wmode, form, delta):
    print(f"[gpu-case] ordered w{wmode} {form} delta={delta}", flush=True)
    g = cg.rmat_graph(12, 16, seed=91)
    s = cg.rmat_stream(g, 12, 6, 400, 400, seed=92)
    og = oracle_graph(g)
    with cg.Engine("sssp", weight_mode=wmode, weight_seed=21, flags=cg.CFG_ORDERED_REPAIR,
                   repair_delta=delta) as e:
        e.load_graph(g)
        e.init_graph(0)
        st = e.start()
        assert st["rounds"] > 0
        check_traversal(e, og, "sssp", 0, wmode, 21, where="ordered initial")
        if form == "grouped":
            e.load_updates(s)
            for lo in (0, 3):
                e.apply_next_batches(3)
                for b in range(lo, lo + 3):
                    oracle_apply(og, *s.batch(b))
                check_traversal(e, og, "sssp", 0, wmode, 21, where=f"ordered group {lo}")
            return
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            e.apply_batch(dels, adds, staged=form == "staged")
            oracle_apply(og, dels, adds)
            check_traversal(e, og, "sssp", 0, wmode, 21, where=f"ordered batch {b}")


@pytest.mark.parametrize("ef,form", [(16, "fused"), (2, "fused"), (2, "grouped"), (4, "host-tier")])
def test_cc_union_repair_matches_oracle(ef, form):
    """Edge factor 2 leaves many components, so deletes split components and
    inserts join them; edge factor 16 has one giant component."""
    print(f"[gpu-case] cc union ef={ef} {form}", flush=True)
    g = cg.rmat_graph(12, ef, seed=93, symmetric=True)
    s = cg.rmat_stream(g, 12, 6, 300, 300, seed=94, symmetric=True)
    og = oracle_graph(g)
    kw = dict(flags=cg.CFG_SYMMETRIC | cg.CFG_CC_UNION_REPAIR)
    if form == "host-tier":
        kw["cache_bytes"] = g.m * 4 // 3
    with cg.Engine("cc", **kw) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        assert og.cc_check(e.values()) == 0, "initial labels"
        assert np.all(e.parents() == cg.INF)
        if form == "grouped":
            e.load_updates(s)
            for lo in (0, 3):
                e.apply_next_batches(3)
                for b in range(lo, lo + 3):
                    oracle_apply(og, *s.batch(b))
                assert og.cc_check(e.values()) == 0, f"group {lo}"
            return
        ncomp = []
        for b in range(s.nbatches):
            dels, adds = s.batch(b)
            st = e.apply_batch(dels, adds)
            oracle_apply(og, dels, adds)
            got = e.values()
            errors = og.cc_check(got)
            assert errors == 0, f"batch {b}: {errors} labels differ"
            ncomp.append(int(np.sum(got == np.arange(g.n))))
            assert st["affected"] > 0
        if ef == 2:
            assert len(set(ncomp)) > 1, "the component count never changed"


def test_cc_union_repair_equals_label_propagation():
    """The same stream through both CC executors gives the same labels."""
    print("[gpu-case] cc union vs propagation", flush=True)
    g = cg.rmat_graph(11, 4, seed=95, symmetric=True)
    s = cg.rmat_stream(g, 11, 5, 200, 200, seed=96, symmetric=True)
    with cg.Engine("cc", flags=cg.CFG_SYMMETRIC | cg.CFG_CC_UNION_REPAIR) as u, \
            cg.Engine("cc", flags=cg.CFG_SYMMETRIC) as p:
        for e in (u, p):
            e.load_graph(g)
            e.init_graph(0)
            e.start()
        assert np.array_equal(u.values(), p.values())
        for b in range(s.nbatches):
            for e in (u, p):
                e.apply_batch(*s.batch(b))
            assert np.array_equal(u.values(), p.values()), b
