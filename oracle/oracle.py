"""ctypes wrapper over liboracle.so — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker (or the timed CPU baseline). The product
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

W_REF128, W_HASH64, W_UNIT, W_ZERO = 0, 1, 2, 3
INF = 0xFFFFFFFF


def build(force: bool = False) -> str:
    """Compile liboracle.so in place with gcc (oracle/Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


class _Graph(C.Structure):
    _fields_ = [("n", C.c_uint32), ("deg", C.c_void_p), ("cap", C.c_void_p), ("rows", C.c_void_p)]


class BellmanResult(C.Structure):
    _fields_ = [("relaxable_edges", C.c_uint64), ("missing_tight_witnesses", C.c_uint64),
                ("first_src", C.c_uint32), ("first_dst", C.c_uint32),
                ("first_missing_witness", C.c_uint32), ("pad", C.c_uint32)]


class ParentResult(C.Structure):
    _fields_ = [("invalid_vertices", C.c_uint64), ("missing_parent_edges", C.c_uint64),
                ("wrong_parent_distances", C.c_uint64), ("first_vertex", C.c_uint32),
                ("first_parent", C.c_uint32)]


class PrCheckResult(C.Structure):
    _fields_ = [("l1_error", C.c_double), ("bound", C.c_double), ("max_error", C.c_double),
                ("invariant_l1", C.c_double), ("max_residual", C.c_double),
                ("reference_mass", C.c_double), ("converged", C.c_int), ("finite", C.c_int),
                ("ok", C.c_int), ("pad", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        L.og_create.restype = C.POINTER(_Graph)
        L.og_create.argtypes = [C.c_uint32, P, P]
        L.og_free.argtypes = [C.POINTER(_Graph)]
        L.og_edge_count.restype = C.c_uint64
        L.og_edge_count.argtypes = [C.POINTER(_Graph)]
        L.og_export.argtypes = [C.POINTER(_Graph), P, P]
        L.og_apply_deletes.argtypes = [C.POINTER(_Graph), P, P, C.c_uint64,
                                       C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.og_apply_adds.argtypes = [C.POINTER(_Graph), P, P, C.c_uint64, C.POINTER(C.c_uint64)]
        for name in ("og_sssp_dijkstra", "og_sssp_label_correcting"):
            getattr(L, name).argtypes = [C.POINTER(_Graph), C.c_uint32, C.c_int, C.c_uint64, P]
        L.og_sssp_label_correcting.restype = C.c_uint64
        L.og_bfs.argtypes = [C.POINTER(_Graph), C.c_uint32, P]
        L.og_cc_flood.argtypes = [C.POINTER(_Graph), P]
        L.og_cc_check.restype = C.c_uint64
        L.og_cc_check.argtypes = [C.POINTER(_Graph), P]
        L.og_pr_jacobi.restype = C.c_int
        L.og_pr_jacobi.argtypes = [C.POINTER(_Graph), P, C.c_double, C.c_int]
        L.og_pr_check.argtypes = [C.POINTER(_Graph), P, P, C.c_double, C.POINTER(PrCheckResult)]
        L.og_apply_batch_mt.restype = C.c_int
        L.og_apply_batch_mt.argtypes = [C.POINTER(_Graph), P, P, C.c_uint64, P, P, C.c_uint64, C.c_uint32,
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.og_pr_invariant.restype = C.c_int
        L.og_pr_invariant.argtypes = [C.POINTER(_Graph), P, P, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.og_check_bellman.argtypes = [C.POINTER(_Graph), P, C.c_uint32, C.c_int, C.c_uint64,
                                       C.POINTER(BellmanResult)]
        L.og_check_parent.argtypes = [C.POINTER(_Graph), P, P, C.c_uint32, C.c_int, C.c_uint64,
                                      C.POINTER(ParentResult)]
        L.og_distance_checksum.restype = C.c_uint64
        L.og_distance_checksum.argtypes = [P, C.c_uint32]
        L.og_parent_checksum.restype = C.c_uint64
        L.og_parent_checksum.argtypes = [P, P, C.c_uint32]
        L.og_weight.restype = C.c_uint32
        L.og_weight.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32]
        L.og_splitmix64.restype = C.c_uint64
        L.og_splitmix64.argtypes = [C.c_uint64]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def weight(mode: int, seed: int, u: int, v: int) -> int:
    return int(lib().og_weight(mode, seed, u, v))


class Graph:
    """Multigraph with the reference's per-source mutation semantics."""

    def __init__(self, n: int, offsets: np.ndarray, dst: np.ndarray):
        self.n = int(n)
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        d = np.ascontiguousarray(dst, dtype=np.uint32)
        if d.size == 0:
            d = np.zeros(1, dtype=np.uint32)
        self._g = lib().og_create(self.n, _p(off), _p(d))
        if not self._g:
            raise MemoryError("og_create failed")

    @classmethod
    def from_edges(cls, n: int, src, dst):
        src = np.asarray(src, dtype=np.uint32)
        dst = np.asarray(dst, dtype=np.uint32)
        order = np.argsort(src, kind="stable")
        counts = np.bincount(src, minlength=n).astype(np.uint64)
        off = np.zeros(n + 1, dtype=np.uint64)
        np.cumsum(counts, out=off[1:])
        return cls(n, off, dst[order])

    def __del__(self):
        g = getattr(self, "_g", None)
        if g:
            lib().og_free(g)
            self._g = None

    def edge_count(self) -> int:
        return int(lib().og_edge_count(self._g))

    def csr(self):
        m = self.edge_count()
        off = np.zeros(self.n + 1, dtype=np.uint64)
        dst = np.zeros(max(m, 1), dtype=np.uint32)
        lib().og_export(self._g, _p(off), _p(dst))
        return off, dst[:m]

    def rows(self):
        off, dst = self.csr()
        return [dst[off[i]:off[i + 1]].tolist() for i in range(self.n)]

    def apply_deletes(self, u, v):
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        miss, inval = C.c_uint64(0), C.c_uint64(0)
        if lib().og_apply_deletes(self._g, _p(u), _p(v), u.size, C.byref(miss), C.byref(inval)):
            raise MemoryError
        return miss.value, inval.value

    def apply_adds(self, u, v):
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        inval = C.c_uint64(0)
        if lib().og_apply_adds(self._g, _p(u), _p(v), u.size, C.byref(inval)):
            raise MemoryError
        return inval.value

    def apply_batch_mt(self, du, dv, au, av, threads: int):
        """Deletes then adds with `threads` workers (the reference's CPU
        mutation pool); returns (missing deletes, invalid records)."""
        arrs = [np.ascontiguousarray(x, dtype=np.uint32) for x in (du, dv, au, av)]
        miss, inval = C.c_uint64(0), C.c_uint64(0)
        if lib().og_apply_batch_mt(self._g, _p(arrs[0]), _p(arrs[1]), arrs[0].size, _p(arrs[2]), _p(arrs[3]),
                                   arrs[2].size, threads, C.byref(miss), C.byref(inval)):
            raise MemoryError
        return miss.value, inval.value

    # ---- recomputes
    def sssp(self, src: int, wmode: int = W_REF128, seed: int = 0) -> np.ndarray:
        out = np.empty(self.n, dtype=np.uint32)
        lib().og_sssp_dijkstra(self._g, src, wmode, seed, _p(out))
        return out

    def sssp_label_correcting(self, src: int, wmode: int = W_REF128, seed: int = 0,
                              with_count: bool = False):
        out = np.empty(self.n, dtype=np.uint32)
        scanned = int(lib().og_sssp_label_correcting(self._g, src, wmode, seed, _p(out)))
        return (out, scanned) if with_count else out

    def bfs(self, src: int) -> np.ndarray:
        out = np.empty(self.n, dtype=np.uint32)
        lib().og_bfs(self._g, src, _p(out))
        return out

    def cc(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.uint32)
        lib().og_cc_flood(self._g, _p(out))
        return out

    def pagerank(self, tol_per_vertex: float = 1e-12, max_rounds: int = 10000):
        x = np.zeros(max(self.n, 1), dtype=np.float64)
        r = lib().og_pr_jacobi(self._g, _p(x), tol_per_vertex, max_rounds)
        return x[: self.n], r

    # ---- checks
    def check_bellman(self, dist, src, wmode=W_REF128, seed=0) -> BellmanResult:
        d = np.ascontiguousarray(dist, dtype=np.uint32)
        res = BellmanResult()
        lib().og_check_bellman(self._g, _p(d), src, wmode, seed, C.byref(res))
        return res

    def check_parent(self, dist, parent, src, wmode=W_REF128, seed=0) -> ParentResult:
        d = np.ascontiguousarray(dist, dtype=np.uint32)
        p = np.ascontiguousarray(parent, dtype=np.uint32)
        res = ParentResult()
        lib().og_check_parent(self._g, _p(d), _p(p), src, wmode, seed, C.byref(res))
        return res

    def cc_check(self, labels) -> int:
        lab = np.ascontiguousarray(labels, dtype=np.uint32)
        return int(lib().og_cc_check(self._g, _p(lab)))

    def pr_check(self, ranks, residual, epsilon) -> PrCheckResult:
        x = np.ascontiguousarray(ranks, dtype=np.float32)
        r = np.ascontiguousarray(residual, dtype=np.float32)
        res = PrCheckResult()
        lib().og_pr_check(self._g, _p(x), _p(r), epsilon, C.byref(res))
        return res

    def pr_invariant(self, ranks, residual):
        """(finite, defect, max|r|, sum|x|): PageRankCheck's invariant
        sum_v |0.15 + 0.85 sum x[u]/deg(u) - x[v] - r[v]| without a reference solve."""
        x = np.ascontiguousarray(ranks, dtype=np.float32)
        r = np.ascontiguousarray(residual, dtype=np.float32)
        d, mr, m = C.c_double(), C.c_double(), C.c_double()
        ok = lib().og_pr_invariant(self._g, _p(x), _p(r), C.byref(d), C.byref(mr), C.byref(m))
        return bool(ok), d.value, mr.value, m.value


def distance_checksum(dist) -> int:
    d = np.ascontiguousarray(dist, dtype=np.uint32)
    return int(lib().og_distance_checksum(_p(d), d.size))


def parent_checksum(dist, parent) -> int:
    d = np.ascontiguousarray(dist, dtype=np.uint32)
    p = np.ascontiguousarray(parent, dtype=np.uint32)
    return int(lib().og_parent_checksum(_p(d), _p(p), d.size))
