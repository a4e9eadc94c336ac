"""cig_query_noalloc: the multi-source query with its working memory in the
library's fixed per-GPU scratch array. Its distances (written to device
memory the test allocates through the library's own HIP runtime) must equal
cig_query's bit for bit on consecutive calls, the second with a smaller input
(fewer sources, a smaller graph), and sizes beyond the array are refused."""
import ctypes as C

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


def _hip():
    """The HIP runtime libcig itself runs on (the one its NEEDED entry loads)."""
    cg.lib()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
    assert paths, "libcig's HIP runtime is not loaded"
    hip = C.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


class DeviceWords:
    """uint32 device buffer owned by the test (the caller's output)."""

    def __init__(self, hip, count):
        self.hip, self.count, self.p = hip, count, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(4, 4 * count)) == 0

    def read(self):
        out = np.empty(self.count, dtype=np.uint32)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 4 * self.count, 2) == 0  # device -> host
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(e, hip, algo, sources):
    k = len(sources)
    buf = DeviceWords(hip, k * e.n)
    try:
        # poison the output so a missing write cannot pass
        poison = np.full(k * e.n, 0xDEADBEEF, dtype=np.uint32)
        assert hip.hipMemcpy(buf.p, poison.ctypes.data, poison.nbytes, 1) == 0  # host -> device
        e.query_noalloc(algo, sources, buf.p.value, 4 * k * e.n)
        return buf.read().reshape(k, e.n)
    finally:
        buf.free()


def test_noalloc_query_matches_query():
    print("[gpu-case] no-alloc query", flush=True)
    hip = _hip()
    g = cg.rmat_graph(12, 16, seed=71)
    s = cg.rmat_stream(g, 12, 1, 300, 300, seed=72)
    og = oracle_graph(g)
    rng = np.random.default_rng(73)
    with cg.Engine("sssp", weight_mode=cg.WEIGHT_HASH64, weight_seed=3) as e:
        e.load_graph(g)
        e.init_graph(0)
        e.start()
        e.apply_batch(*s.batch(0))
        oracle_apply(og, *s.batch(0))
        before = e.values().copy()
        src64 = rng.integers(0, g.n, 64)
        src64[0] = 0
        for algo in ("bfs", "sssp"):
            got = _noalloc(e, hip, algo, src64)  # first call: 64 sources
            assert np.array_equal(got, e.query(algo, src64)), algo
            for j in (0, 17, 63):
                exp = og.bfs(int(src64[j])) if algo == "bfs" else og.sssp(int(src64[j]), O.W_HASH64, 3)
                assert np.array_equal(got[j], exp), (algo, j)
        src7 = rng.integers(0, g.n, 7)
        got = _noalloc(e, hip, "sssp", src7)  # second call: fewer sources
        assert np.array_equal(got, e.query("sssp", src7))
        assert np.array_equal(e.values(), before), "the query changed the engine state"
    # a smaller graph on another engine: the array still holds the last call's contents
    g2 = cg.rmat_graph(9, 8, seed=74)
    with cg.Engine("bfs") as e2:
        e2.load_graph(g2)
        e2.init_graph(0)
        e2.start()
        src = np.array([0, 5, 9], dtype=np.uint32)
        for algo in ("sssp", "bfs"):
            got = _noalloc(e2, hip, algo, src)
            assert np.array_equal(got, e2.query(algo, src)), algo


def test_noalloc_query_rejects_sizes_beyond_the_array():
    print("[gpu-case] no-alloc limit", flush=True)
    hip = _hip()
    g = cg.rmat_graph(20, 1, seed=75)  # 2^20 vertices: 64 sources need ~298 MB > 256 MiB
    with cg.Engine("sssp") as e:
        e.load_graph(g)
        n = e.n
        need64 = n * (4 * 64 + 28)
        assert need64 > cg.NOALLOC_SCRATCH_BYTES
        buf = DeviceWords(hip, 64 * n)
        try:
            with pytest.raises(cg.CigError) as ei:
                e.query_noalloc("sssp", np.arange(64), buf.p.value, 4 * 64 * n)
            assert ei.value.status == cg.ELIMIT
        finally:
            buf.free()
        src = np.array([0, 1, 2, 3, 4, 5, 6, 7], dtype=np.uint32)  # fits: 60 B per vertex
        got = _noalloc(e, hip, "sssp", src)
        assert np.array_equal(got, e.query("sssp", src))
