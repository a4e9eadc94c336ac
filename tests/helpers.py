"""Shared test helpers: run the HIP engine and the CPU oracle in lockstep."""
from __future__ import annotations

import numpy as np

import coincgraph_amd as cg
from oracle import oracle as O

ORACLE_WMODE = {cg.WEIGHT_REF128: O.W_REF128, cg.WEIGHT_HASH64: O.W_HASH64}


def oracle_graph(g: cg.HostGraph) -> O.Graph:
    return O.Graph(g.n, g.offsets, g.dst)


def oracle_values(og: O.Graph, algo: str, source=0, wmode=cg.WEIGHT_REF128, seed=0):
    if algo == "sssp":
        return og.sssp(source, ORACLE_WMODE[wmode], seed)
    if algo == "bfs":
        return og.bfs(source)
    if algo == "cc":
        return og.cc()
    raise ValueError(algo)


def oracle_apply(og: O.Graph, dels, adds):
    miss, inval = og.apply_deletes(dels["u"], dels["v"])
    inval += og.apply_adds(adds["u"], adds["v"])
    return miss, inval


def check_traversal(eng: cg.Engine, og: O.Graph, algo: str, source=0, wmode=cg.WEIGHT_REF128,
                    seed=0, where=""):
    """Bit-exact values + valid parent witnesses (the reference's --check)."""
    got = eng.values()
    exp = oracle_values(og, algo, source, wmode, seed)
    bad = np.nonzero(got != exp)[0]
    assert bad.size == 0, (f"{where}: {bad.size} mismatches, first v={bad[:5].tolist()} "
                           f"got={got[bad[:5]].tolist()} exp={exp[bad[:5]].tolist()}")
    if algo in ("sssp", "bfs"):
        om = ORACLE_WMODE[wmode] if algo == "sssp" else O.W_UNIT
        par = eng.parents()
        res = og.check_parent(got, par, source, om, seed)
        assert res.invalid_vertices == 0, f"{where}: parent witness invalid at v={res.first_vertex}"
        bel = og.check_bellman(got, source, om, seed)
        assert bel.relaxable_edges == 0 and bel.missing_tight_witnesses == 0, where
    elif algo == "cc":
        assert og.cc_check(got) == 0, where
        # parent witnesses for labels: an in-neighbour with the same label
        par = eng.parents()
        rows = None
        for v in np.nonzero(got != np.arange(len(got)))[0][:200]:
            p = int(par[v])
            assert p != cg.INF and got[p] == got[v], f"{where}: CC parent of {v}"
            if rows is None:
                off, dst = og.csr()
            assert v in dst[off[p]:off[p + 1]], f"{where}: CC parent edge ({p},{v}) missing"


def check_topology(eng: cg.Engine, og: O.Graph, where=""):
    """Every row holds the same multiset of destinations as the oracle's."""
    t = eng.topology()
    off, dst = og.csr()
    assert np.array_equal(np.diff(t.offsets), np.diff(off)), f"{where}: degrees differ"
    a = np.lexsort((t.dst, np.repeat(np.arange(t.n), np.diff(t.offsets).astype(np.int64))))
    b = np.lexsort((dst, np.repeat(np.arange(og.n), np.diff(off).astype(np.int64))))
    assert np.array_equal(t.dst[a], dst[b]), f"{where}: row multisets differ"


def check_incoming(eng: cg.Engine, og: O.Graph, where=""):
    """The device incoming-row index: row v holds the multiset of u over the
    live edges (u, v), one entry per parallel edge."""
    t = eng.incoming()
    off, dst = og.csr()
    src = np.repeat(np.arange(og.n), np.diff(off).astype(np.int64))
    assert np.array_equal(np.diff(t.offsets).astype(np.int64), np.bincount(dst, minlength=og.n)), \
        f"{where}: in-degrees differ"
    exp = src[np.lexsort((src, dst))]
    rows = np.repeat(np.arange(t.n), np.diff(t.offsets).astype(np.int64))
    assert np.array_equal(t.dst[np.lexsort((t.dst, rows))], exp), f"{where}: incoming rows differ"


def pr_errors(ranks: np.ndarray, ref: np.ndarray):
    l1 = float(np.abs(ranks.astype(np.float64) - ref).sum())
    mass = float(np.abs(ref).sum())
    return l1 / max(mass, 1e-300), float(np.abs(ranks.astype(np.float64) - ref).max())
