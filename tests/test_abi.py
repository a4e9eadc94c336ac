"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, and its host utilities (R-MAT generator, update stream,
readers) behave as documented. No compute call needs a GPU here."""
import ctypes as C
import os

import numpy as np
import pytest

import coincgraph_amd as cg
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_library_exports_every_header_symbol():
    L = cg.lib()
    names = cg.header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name


def test_config_defaults_and_create_errors():
    cfg = cg.Config()
    assert cg.lib().cig_config_default(C.byref(cfg)) == cg.OK
    assert cfg.abi_version == cg.ABI_VERSION and cfg.algo == cg.SSSP
    assert cfg.pr_max_rounds == 100 and abs(cfg.pr_epsilon - 1e-6) < 1e-12
    bad = cg.Config()
    cg.lib().cig_config_default(C.byref(bad))
    bad.abi_version = 99
    e = C.c_void_p()
    assert cg.lib().cig_create(C.byref(bad), C.byref(e)) == cg.EINVAL
    bad.abi_version = cg.ABI_VERSION
    bad.algo = cg.PR
    bad.pr_max_rounds = 0  # at least one round (hybrid_pr.cu:129-132)
    assert cg.lib().cig_create(C.byref(bad), C.byref(e)) == cg.EINVAL
    assert b"pr_max_rounds" in cg.lib().cig_last_error(None)


def test_create_without_device_fails_cleanly():
    if cg.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(cg.CigError) as err:
        cg.Engine("sssp")
    assert err.value.status == cg.EDEVICE


def test_null_engine_calls_are_rejected():
    L = cg.lib()
    assert L.cig_converge(None, None) == cg.EINVAL
    assert L.cig_gather(None, 0, None, 0) == cg.EINVAL
    assert L.cig_update_batches_left(None) == 0


def test_edge_weight_matches_oracle():
    for u, v in [(0, 0), (1, 2), (127, 1), (0xFFFFFFF0, 77), (123456, 654321)]:
        assert cg.edge_weight(cg.WEIGHT_REF128, 0, u, v) == O.weight(O.W_REF128, 0, u, v)
        assert cg.edge_weight(cg.WEIGHT_HASH64, 42, u, v) == O.weight(O.W_HASH64, 42, u, v)


def test_rmat_generator_properties():
    g1 = cg.rmat_graph(12, 16, seed=7, threads=1)
    g8 = cg.rmat_graph(12, 16, seed=7, threads=8)
    assert np.array_equal(g1.offsets, g8.offsets) and np.array_equal(g1.dst, g8.dst)
    g2 = cg.rmat_graph(12, 16, seed=8)
    assert not np.array_equal(g1.dst[:100], g2.dst[:100])
    n = g1.n
    assert n == 4096 and g1.offsets[-1] == g1.dst.size
    src = np.repeat(np.arange(n), np.diff(g1.offsets).astype(np.int64))
    assert np.all(src != g1.dst), "self-loops must be removed"
    key = src.astype(np.uint64) << np.uint64(32) | g1.dst.astype(np.uint64)
    assert np.unique(key).size == key.size, "duplicates must be removed"
    deg = np.diff(g1.offsets)
    assert deg.argmax() == 0, "unpermuted R-MAT: vertex 0 is the hub"
    # quadrant mass: share of edges with both ids in the low half ~ a^1 at depth 1
    low = (src < n // 2) & (g1.dst < n // 2)
    assert 0.5 < low.mean() < 0.75
    gs = cg.rmat_graph(10, 8, seed=3, symmetric=True)
    ssrc = np.repeat(np.arange(gs.n), np.diff(gs.offsets).astype(np.int64))
    fwd = set(zip(ssrc.tolist(), gs.dst.tolist()))
    assert all((v, u) in fwd for u, v in fwd)


def test_update_stream_properties():
    g = cg.rmat_graph(11, 8, seed=4)
    s = cg.rmat_stream(g, 11, 5, 200, 150, seed=5)
    assert s.batch_sizes.tolist() == [[150, 200]] * 5
    og = O.Graph(g.n, g.offsets, g.dst)
    for b in range(5):
        dels, adds = s.batch(b)
        rows = og.rows()
        present = {(u, v) for u in range(g.n) for v in rows[u]}
        assert all((int(u), int(v)) in present for u, v in zip(dels["u"], dels["v"]))
        keys = set(zip(dels["u"].tolist(), dels["v"].tolist()))
        assert len(keys) == len(dels), "deletes sampled without replacement"
        miss, inval = og.apply_deletes(dels["u"], dels["v"])
        assert miss == 0 and inval == 0
        rows = og.rows()
        present = {(u, v) for u in range(g.n) for v in rows[u]}
        assert not any((int(u), int(v)) in present for u, v in zip(adds["u"], adds["v"]))
        assert np.all(adds["u"] != adds["v"])
        og.apply_adds(adds["u"], adds["v"])
    ss = cg.rmat_stream(cg.rmat_graph(10, 8, seed=6, symmetric=True), 10, 2, 50, 60, seed=7,
                        symmetric=True)
    d, a = ss.batch(0)
    assert len(d) == 100 and len(a) == 120
    assert set(zip(d["u"].tolist(), d["v"].tolist())) == set(zip(d["v"].tolist(), d["u"].tolist()))


def test_readers_on_golden_fixture():
    g = cg.read_market(os.path.join(GOLDEN, "readme", "graph.txt"))
    assert g.n == 5 and g.offsets.tolist() == [0, 2, 3, 4, 5, 5]
    assert g.dst.tolist() == [1, 2, 3, 3, 4]
    s = cg.read_updates(os.path.join(GOLDEN, "readme", "updates.txt"),
                        os.path.join(GOLDEN, "readme", "batches.txt"))
    assert s.nbatches == 1 and s.batch_sizes.tolist() == [[1, 1]]
    assert (s.dels["u"][0], s.dels["v"][0]) == (1, 3)
    assert (s.adds["u"][0], s.adds["v"][0]) == (2, 1)


def test_readers_reject_bad_files(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("0 1\nx\n")
    with pytest.raises(cg.CigError):
        cg.read_market(str(p))
    with pytest.raises(cg.CigError):
        cg.read_market(str(tmp_path / "missing.txt"))
    u = tmp_path / "u.txt"
    u.write_text("a 0 1 1\n")
    bs = tmp_path / "b.txt"
    bs.write_text("2 0\n")  # asks for more records than the file holds
    with pytest.raises(cg.CigError):
        cg.read_updates(str(u), str(bs))


def test_binary_formats_round_trip(tmp_path):
    """Binary graph and stream files hold exactly what the text readers and
    the generators produce; the golden README fixture survives text -> bin."""
    g = cg.read_market(os.path.join(GOLDEN, "readme", "graph.txt"))
    s = cg.read_updates(os.path.join(GOLDEN, "readme", "updates.txt"),
                        os.path.join(GOLDEN, "readme", "batches.txt"))
    big = cg.rmat_graph(10, 8, seed=21)
    bs = cg.rmat_stream(big, 10, 3, 40, 50, seed=22)
    for name, graph, stream in (("readme", g, s), ("rmat", big, bs)):
        gp, sp = str(tmp_path / f"{name}.cigg"), str(tmp_path / f"{name}.cigu")
        cg.write_csr_bin(graph, gp)
        cg.write_updates_bin(stream, sp)
        g2, s2 = cg.read_csr_bin(gp), cg.read_updates_bin(sp)
        assert g2.n == graph.n
        assert np.array_equal(g2.offsets, graph.offsets) and np.array_equal(g2.dst, graph.dst)
        assert np.array_equal(s2.batch_sizes, stream.batch_sizes)
        assert np.array_equal(s2.adds, stream.adds) and np.array_equal(s2.dels, stream.dels)
        # header (32 B) + arrays, nothing else
        assert os.path.getsize(gp) == 32 + 8 * (graph.n + 1) + 4 * graph.m
        assert os.path.getsize(sp) == 32 + 8 * stream.nbatches + 12 * (len(stream.adds) + len(stream.dels))


def test_binary_readers_reject_bad_files(tmp_path):
    g = cg.rmat_graph(8, 4, seed=23)
    s = cg.rmat_stream(g, 8, 2, 10, 10, seed=24)
    gp, sp = tmp_path / "g.cigg", tmp_path / "s.cigu"
    cg.write_csr_bin(g, str(gp))
    cg.write_updates_bin(s, str(sp))
    raw_g, raw_s = gp.read_bytes(), sp.read_bytes()
    cases = {
        "g_trunc.cigg": raw_g[:-4],
        "g_extra.cigg": raw_g + b"\0",
        "g_magic.cigg": b"XXXX" + raw_g[4:],
        "s_trunc.cigu": raw_s[:-1],
        "s_magic.cigu": raw_g,  # a graph file is not a stream
    }
    for name, data in cases.items():
        (tmp_path / name).write_bytes(data)
        with pytest.raises(cg.CigError):
            (cg.read_csr_bin if name.endswith(".cigg") else cg.read_updates_bin)(str(tmp_path / name))
    # batch sizes asking for more records than the file holds
    big_sizes = bytearray(raw_s)
    big_sizes[32:36] = (10**6).to_bytes(4, "little")
    (tmp_path / "s_sizes.cigu").write_bytes(bytes(big_sizes))
    with pytest.raises(cg.CigError):
        cg.read_updates_bin(str(tmp_path / "s_sizes.cigu"))
    # header record counts whose byte size wraps 2^64 (2^61 * 8 = 0) or exceeds the
    # file are rejected as I/O errors before anything is allocated
    for off, count in ((16, 1 << 61), (16, (1 << 64) - 1), (24, 1 << 62)):
        for raw, ext in ((raw_g, "cigg"), (raw_s, "cigu")):
            if off == 24 and ext == "cigg":
                continue  # the graph header's second count is unused
            bad = bytearray(raw)
            bad[off:off + 8] = count.to_bytes(8, "little")
            p = tmp_path / f"count_{off}_{count}.{ext}"
            p.write_bytes(bytes(bad))
            with pytest.raises(cg.CigError) as ei:
                (cg.read_csr_bin if ext == "cigg" else cg.read_updates_bin)(str(p))
            assert ei.value.status == cg.EIO, (off, count, ext)


def test_oracle_golden_fixture_known_answers():
    """The oracle reproduces the hand-derived README fixture answers."""
    d = os.path.join(GOLDEN, "readme")
    e = np.loadtxt(os.path.join(d, "graph.txt"), dtype=np.int64, ndmin=2)
    og = O.Graph.from_edges(5, e[:, 0], e[:, 1])
    og.apply_deletes([1], [3])
    og.apply_adds([2], [1])
    exp = {k: np.loadtxt(os.path.join(d, f"expected_{k}.txt"), ndmin=2)[:, 1]
           for k in ("sssp", "bfs", "cc", "pr")}
    assert og.sssp(0).tolist() == exp["sssp"].astype(int).tolist()
    assert og.bfs(0).tolist() == exp["bfs"].astype(int).tolist()
    assert og.cc().tolist() == exp["cc"].astype(int).tolist()
    x, _ = og.pagerank()
    np.testing.assert_allclose(x, exp["pr"], rtol=0, atol=1e-12)


def test_alternative_executor_configs_are_checked():
    """CIG_CFG_CC_UNION_REPAIR needs CC on an undirected graph; refused at
    cig_create, before any device is touched."""
    cfg = cg.Config()
    L = cg.lib()
    e = C.c_void_p()
    for algo, flags in ((cg.CC, cg.CFG_CC_UNION_REPAIR), (cg.SSSP, cg.CFG_SYMMETRIC | cg.CFG_CC_UNION_REPAIR),
                        (cg.CC, cg.CFG_SYMMETRIC | cg.CFG_CC_UNION_REPAIR | cg.CFG_COOPERATIVE)):
        L.cig_config_default(C.byref(cfg))
        cfg.algo, cfg.flags = algo, flags
        assert L.cig_create(C.byref(cfg), C.byref(e)) == cg.EINVAL
        assert b"UNION_REPAIR" in L.cig_last_error(None)


def test_out_of_core_generator_matches_in_memory(tmp_path):
    """cig_rmat_write (source segments, bounded passes) writes exactly the graph
    cig_rmat_csr builds from the same arguments, directed and undirected, for
    one pass and for many."""
    for scale, ef, sym, segments, mem in ((11, 8, False, 16, 1 << 12), (11, 8, True, 8, 1 << 14),
                                          (10, 4, False, 1, 1 << 30), (12, 16, False, 512, 1 << 16)):
        ref = cg.rmat_graph(scale, ef, seed=scale * 7, symmetric=sym)
        path = str(tmp_path / f"g{scale}_{ef}_{sym}_{segments}.cigg")
        m = cg.rmat_write(path, scale, ef, seed=scale * 7, symmetric=sym, segments=segments, mem_bytes=mem)
        got = cg.read_csr_bin(path)
        assert m == ref.m and got.n == ref.n
        assert np.array_equal(got.offsets, ref.offsets), (scale, sym, segments)
        assert np.array_equal(got.dst, ref.dst), (scale, sym, segments)
    with pytest.raises(cg.CigError):
        cg.rmat_write(str(tmp_path / "bad.cigg"), 10, 4, segments=3)  # not a power of two
