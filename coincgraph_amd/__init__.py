"""coincgraph_amd: MI355X incremental graph engine (BFS / SSSP / CC / PageRank).

The compute lives in libcig.so (hand-written HIP kernels for gfx950 behind the
C-ABI of include/cig.h). This module is a thin ctypes binding plus an
`Engine` class whose methods mirror the reference engine the samples drive
(sepgraph::engine::Engine, include/framework/framework.cuh:456-3505):
LoadGraph, InitGraph, Start, del_edge, add_edge, UpdatePageRank,
cache maintenance and the Gather* result copies.

There is no CPU fallback: every call goes to the HIP library, and a missing
library or GPU raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import re
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcig.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "cig.h")

# cig_status
OK, EINVAL, ENOMEM, EDEVICE, ESTATE, EIO, ELIMIT, EINTERNAL, EPENDING = range(9)
STATUS_NAMES = ["OK", "EINVAL", "ENOMEM", "EDEVICE", "ESTATE", "EIO", "ELIMIT", "EINTERNAL", "EPENDING"]
# cig_algo
BFS, SSSP, CC, PR = range(4)
ALGOS = {"bfs": BFS, "sssp": SSSP, "cc": CC, "pr": PR}
# cig_weight_mode
WEIGHT_REF128, WEIGHT_HASH64 = 0, 1
# cig_field
FIELD_VALUE, FIELD_PARENT, FIELD_RESIDUAL, FIELD_DEGREE, FIELD_TIER, FIELD_HOTNESS = range(6)
# flags
CFG_COOPERATIVE, CFG_NO_HOTNESS = 1, 2
CFG_SCAN_REPAIR = 1 << 2
CFG_KERNEL_TIMING = 1 << 3
CFG_PR_PUSH_ONLY = 1 << 4
CFG_SYMMETRIC = 1 << 5
CFG_ORDERED_REPAIR = 1 << 6
CFG_CC_UNION_REPAIR = 1 << 7
CFG_GRAPH_BATCHES = 1 << 8
CFG_PR_KEEP_IDS = 1 << 9
BATCH_STAGED = 1
BATCHES_EACH = 1 << 3  # CIG_BATCHES_EACH
INF = 0xFFFFFFFF
ABI_VERSION = 1
NOALLOC_SCRATCH_BYTES = 256 << 20  # CIG_NOALLOC_SCRATCH_BYTES


class CigError(RuntimeError):
    def __init__(self, status: int, message: str):
        name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"{name}: {message}")
        self.status = status


class Edge(C.Structure):
    _fields_ = [("u", C.c_uint32), ("v", C.c_uint32), ("w", C.c_uint32)]


EDGE_DTYPE = np.dtype([("u", "<u4"), ("v", "<u4"), ("w", "<u4")])


class Config(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("algo", C.c_uint32), ("weight_mode", C.c_uint32),
        ("flags", C.c_uint32), ("weight_seed", C.c_uint64), ("cache_bytes", C.c_uint64),
        ("device", C.c_int32), ("n_stream", C.c_uint32), ("segments", C.c_uint32),
        ("block_size", C.c_uint32), ("edges_per_lane", C.c_uint32),
        ("split_degree", C.c_uint32), ("max_rounds", C.c_uint32),
        ("pr_max_rounds", C.c_uint32), ("pr_epsilon", C.c_float),
        ("tile_len", C.c_uint32), ("round_blocks_per_cu", C.c_uint32),
        ("pr_pull_divisor", C.c_uint32), ("pull_blocks_per_cu", C.c_uint32),
        ("kernel_timing_every", C.c_uint32), ("repair_delta", C.c_uint32), ("pool_slack", C.c_uint32),
        ("adapt_split", C.c_uint32),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double), ("mutate_ms", C.c_double), ("invalidate_ms", C.c_double),
        ("repair_ms", C.c_double), ("seed_ms", C.c_double), ("propagate_ms", C.c_double),
        ("cache_ms", C.c_double),
        ("edges_scanned", C.c_uint64), ("host_edges_scanned", C.c_uint64),
        ("vertices_expanded", C.c_uint64), ("affected", C.c_uint64), ("rounds", C.c_uint64),
        ("deletes_applied", C.c_uint64), ("missing_deletes", C.c_uint64),
        ("invalid_updates", C.c_uint64), ("adds_applied", C.c_uint64),
        ("relocated_rows", C.c_uint64), ("pool_compactions", C.c_uint64),
        ("touched_sources", C.c_uint64), ("converged", C.c_uint32), ("launches", C.c_uint32),
        ("pr_residual_l1", C.c_double), ("stage_edges", C.c_uint64 * 4),
        ("stage_vertices", C.c_uint64 * 4), ("stage_host_edges", C.c_uint64 * 4),
        ("scan_kernel_ms", C.c_double), ("round_kernel_ms", C.c_double),
        ("round_launches", C.c_uint64), ("pull_rounds", C.c_uint64), ("inval_rounds", C.c_uint64),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            if name == "reserved":
                continue
            v = getattr(self, name)
            out[name] = list(v) if name.startswith("stage_") else v
        return out


class TuneResult(C.Structure):
    _fields_ = [("edges_per_lane", C.c_uint32), ("split_degree", C.c_uint32), ("tile_len", C.c_uint32),
                ("block_size", C.c_uint32), ("round_blocks_per_cu", C.c_uint32), ("trials", C.c_uint32),
                ("first_ms", C.c_double), ("best_ms", C.c_double), ("reserved", C.c_uint64 * 2)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class ServeConfig(C.Structure):
    _fields_ = [("max_group", C.c_uint32), ("max_query_sources", C.c_uint32),
                ("hbm_reserve", C.c_uint64), ("refresh_every", C.c_uint32),
                ("max_query_wait_ms", C.c_uint32), ("reserved", C.c_uint64 * 3)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_uint32), ("kind", C.c_uint32), ("version", C.c_uint64),
                ("submit_ms", C.c_double), ("start_ms", C.c_double), ("done_ms", C.c_double),
                ("group_size", C.c_uint32), ("reserved0", C.c_uint32), ("stats", BatchStats)]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_ if name not in ("reserved0", "stats")}
        d["stats"] = self.stats.as_dict()
        return d


class GraphInfo(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("algo", C.c_uint32), ("edges", C.c_uint64),
        ("hbm_pool_slots", C.c_uint64), ("hbm_pool_used", C.c_uint64),
        ("host_pool_slots", C.c_uint64), ("host_pool_used", C.c_uint64),
        ("cached_rows", C.c_uint64), ("cached_edges", C.c_uint64),
        ("device_bytes", C.c_uint64), ("host_pinned_bytes", C.c_uint64),
        ("grid_blocks", C.c_uint32), ("block_size", C.c_uint32),
        ("edges_per_lane", C.c_uint32), ("split_degree", C.c_uint32),
        ("cu_count", C.c_uint32), ("compactions", C.c_uint32), ("inplace_compactions", C.c_uint32),
        ("host_compactions", C.c_uint32), ("compaction_ms_hbm", C.c_uint32),
        ("compaction_ms_host", C.c_uint32), ("refreshes_deferred", C.c_uint32), ("reserved", C.c_uint32 * 1),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


_lib = None
# Engines and servers still open. They are closed by an interpreter exit hook,
# which runs before the C runtime's exit handlers (and the HIP runtime's
# teardown among them): no HIP call is left to a finaliser that runs after
# the runtime is gone, and a server thread never outlives its engine.
_LIVE = weakref.WeakSet()


@atexit.register
def _close_all():
    objs = list(_LIVE)
    for o in objs:
        if isinstance(o, Server):
            try:
                o.stop()
            except Exception:
                pass
    for o in objs:
        if isinstance(o, Engine):
            o.close()


def header_functions(path: str = HEADER_PATH) -> list:
    """Names of the functions include/cig.h declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cig_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Loads libcig.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CigError(EDEVICE, f"{LIB_PATH} missing: run python -m coincgraph_amd.build")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    eng = C.c_void_p
    sig = {
        "cig_config_default": (C.c_int, [C.POINTER(Config)]),
        "cig_create": (C.c_int, [C.POINTER(Config), C.POINTER(eng)]),
        "cig_destroy": (None, [eng]),
        "cig_last_error": (C.c_char_p, [eng]),
        "cig_load_csr": (C.c_int, [eng, C.c_uint32, P, P]),
        "cig_graph_info_get": (C.c_int, [eng, C.POINTER(GraphInfo)]),
        "cig_gather_topology": (C.c_int, [eng, P, P, C.c_uint64]),
        "cig_gather_incoming": (C.c_int, [eng, P, P, C.c_uint64]),
        "cig_init": (C.c_int, [eng, P, C.c_uint32]),
        "cig_converge": (C.c_int, [eng, C.POINTER(BatchStats)]),
        "cig_delete_edges": (C.c_int, [eng, P, C.c_uint64, C.POINTER(BatchStats)]),
        "cig_insert_edges": (C.c_int, [eng, P, C.c_uint64, C.POINTER(BatchStats)]),
        "cig_apply_batch": (C.c_int, [eng, P, C.c_uint64, P, C.c_uint64, C.c_uint32,
                                      C.POINTER(BatchStats)]),
        "cig_load_updates": (C.c_int, [eng, P, C.c_uint64, P, C.c_uint64, P, C.c_uint32]),
        "cig_apply_next_batch": (C.c_int, [eng, C.c_uint32, C.POINTER(BatchStats)]),
        "cig_apply_next_batches": (C.c_int, [eng, C.c_uint32, C.c_uint32, C.POINTER(BatchStats)]),
        "cig_query": (C.c_int, [eng, C.c_uint32, P, C.c_uint32, P, C.c_uint64, C.POINTER(BatchStats)]),
        "cig_query_noalloc": (C.c_int, [eng, C.c_uint32, P, C.c_uint32, C.c_void_p, C.c_uint64,
                                        C.POINTER(BatchStats)]),
        "cig_update_batches_left": (C.c_uint32, [eng]),
        "cig_cache_refresh": (C.c_int, [eng, C.POINTER(BatchStats)]),
        "cig_compact": (C.c_int, [eng, C.c_uint32]),
        "cig_gather": (C.c_int, [eng, C.c_uint32, P, C.c_size_t]),
        "cig_rmat_csr": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                   C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32),
                                   C.POINTER(P), C.POINTER(P)]),
        "cig_rmat_stream": (C.c_int, [C.c_uint32, P, P, C.c_uint32, C.c_double, C.c_double,
                                      C.c_double, C.c_uint64, C.c_uint32, C.c_uint64,
                                      C.c_uint64, C.c_uint32, C.POINTER(P),
                                      C.POINTER(C.c_uint64), C.POINTER(P),
                                      C.POINTER(C.c_uint64), C.POINTER(P)]),
        "cig_read_market": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(P),
                                      C.POINTER(P)]),
        "cig_read_updates": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(P),
                                       C.POINTER(C.c_uint64), C.POINTER(P),
                                       C.POINTER(C.c_uint64), C.POINTER(P),
                                       C.POINTER(C.c_uint32)]),
        "cig_write_csr_bin": (C.c_int, [C.c_char_p, C.c_uint32, P, P]),
        "cig_read_csr_bin": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(P), C.POINTER(P)]),
        "cig_rmat_write": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                     C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                     C.POINTER(C.c_uint64)]),
        "cig_write_updates_bin": (C.c_int, [C.c_char_p, P, C.c_uint64, P, C.c_uint64, P, C.c_uint32]),
        "cig_read_updates_bin": (C.c_int, [C.c_char_p, C.POINTER(P), C.POINTER(C.c_uint64), C.POINTER(P),
                                           C.POINTER(C.c_uint64), C.POINTER(P), C.POINTER(C.c_uint32)]),
        "cig_host_free": (None, [P]),
        "cig_edge_weight": (C.c_uint32, [C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
        "cig_device_count": (C.c_int, []),
        "cig_copy_probe": (C.c_int, [C.c_int32, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
        "cig_host_read_probe": (C.c_int, [C.c_int32, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
        "cig_autotune": (C.c_int, [eng, C.c_uint32, C.POINTER(TuneResult)]),
        "cig_serve_config_default": (C.c_int, [C.POINTER(ServeConfig)]),
        "cig_serve_start": (C.c_int, [eng, C.POINTER(ServeConfig), C.POINTER(P)]),
        "cig_submit_update": (C.c_int, [P, P, C.c_uint64, P, C.c_uint64, C.POINTER(C.c_uint64)]),
        "cig_submit_query": (C.c_int, [P, C.c_uint32, P, C.c_uint32, P, C.c_uint64,
                                       C.POINTER(C.c_uint64)]),
        "cig_poll": (C.c_int, [P, C.c_uint64, C.POINTER(RequestInfo)]),
        "cig_wait": (C.c_int, [P, C.c_uint64, C.POINTER(RequestInfo)]),
        "cig_serve_stop": (C.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _take(p, count, dtype):
    """Copies `count` elements from a library-malloc'd buffer and frees it."""
    if not p.value:
        return np.zeros(0, dtype=dtype)
    nbytes = int(count) * np.dtype(dtype).itemsize
    out = np.empty(count, dtype=dtype)
    if nbytes:
        C.memmove(out.ctypes.data, p.value, nbytes)
    lib().cig_host_free(p)
    return out


def _check(status):
    if status != OK:
        raise CigError(status, "host utility failed")


def edges_array(u, v, w=None) -> np.ndarray:
    u = np.asarray(u, dtype=np.uint32)
    v = np.asarray(v, dtype=np.uint32)
    a = np.empty(len(u), dtype=EDGE_DTYPE)
    a["u"] = u
    a["v"] = v
    a["w"] = 1 if w is None else np.asarray(w, dtype=np.uint32)
    return a


def device_count() -> int:
    return int(lib().cig_device_count())


def copy_probe(bytes_: int, iters: int = 20, device: int = 0) -> float:
    """Mean milliseconds of a plain copy kernel moving `bytes_` bytes."""
    ms = C.c_double()
    _check(lib().cig_copy_probe(device, bytes_, iters, C.byref(ms)))
    return ms.value


def host_read_probe(bytes_: int, iters: int = 10, device: int = 0) -> float:
    """Mean milliseconds of a kernel reading `bytes_` bytes of pinned host memory zero-copy."""
    ms = C.c_double()
    _check(lib().cig_host_read_probe(device, bytes_, iters, C.byref(ms)))
    return ms.value


def edge_weight(mode: int, seed: int, u: int, v: int) -> int:
    return int(lib().cig_edge_weight(mode, seed, u, v))


# ---------------------------------------------------------------- host utils
@dataclass
class HostGraph:
    n: int
    offsets: np.ndarray  # uint64, n+1
    dst: np.ndarray      # uint32

    @property
    def m(self) -> int:
        return int(self.offsets[-1])


@dataclass
class UpdateStream:
    adds: np.ndarray        # EDGE_DTYPE
    dels: np.ndarray        # EDGE_DTYPE
    batch_sizes: np.ndarray  # uint32 (nbatches, 2): adds, dels

    @property
    def nbatches(self) -> int:
        return int(self.batch_sizes.shape[0])

    def batch(self, b: int):
        sizes = self.batch_sizes
        a0 = int(sizes[:b, 0].sum())
        d0 = int(sizes[:b, 1].sum())
        return (self.dels[d0:d0 + int(sizes[b, 1])], self.adds[a0:a0 + int(sizes[b, 0])])


RMAT_ABC = (0.57, 0.19, 0.19)


def rmat_graph(scale: int, edge_factor: int = 16, seed: int = 1, symmetric: bool = False,
               abc=RMAT_ABC, threads: int = 0) -> HostGraph:
    n = C.c_uint32()
    off = C.c_void_p()
    dst = C.c_void_p()
    _check(lib().cig_rmat_csr(scale, edge_factor, abc[0], abc[1], abc[2], seed,
                              1 if symmetric else 0, threads, C.byref(n), C.byref(off),
                              C.byref(dst)))
    offsets = _take(off, n.value + 1, np.uint64)
    return HostGraph(n.value, offsets, _take(dst, int(offsets[-1]), np.uint32))


def rmat_stream(g: HostGraph, scale: int, nbatches: int, n_del: int, n_add: int, seed: int = 2,
                symmetric: bool = False, abc=RMAT_ABC) -> UpdateStream:
    adds = C.c_void_p()
    dels = C.c_void_p()
    sizes = C.c_void_p()
    na = C.c_uint64()
    nd = C.c_uint64()
    _check(lib().cig_rmat_stream(g.n, _ptr(g.offsets), _ptr(g.dst) if g.m else None, scale,
                                 abc[0], abc[1], abc[2], seed, nbatches, n_del, n_add,
                                 1 if symmetric else 0, C.byref(adds), C.byref(na),
                                 C.byref(dels), C.byref(nd), C.byref(sizes)))
    return UpdateStream(_take(adds, na.value, EDGE_DTYPE), _take(dels, nd.value, EDGE_DTYPE),
                        _take(sizes, 2 * nbatches, np.uint32).reshape(nbatches, 2))


def read_market(path: str) -> HostGraph:
    n = C.c_uint32()
    off = C.c_void_p()
    dst = C.c_void_p()
    _check(lib().cig_read_market(path.encode(), C.byref(n), C.byref(off), C.byref(dst)))
    offsets = _take(off, n.value + 1, np.uint64)
    return HostGraph(n.value, offsets, _take(dst, int(offsets[-1]), np.uint32))


def write_csr_bin(g: HostGraph, path: str) -> None:
    """Binary graph file ("CIGG", see cig.h)."""
    _check(lib().cig_write_csr_bin(path.encode(), g.n, _ptr(g.offsets), _ptr(g.dst) if g.m else None))


def rmat_write(path: str, scale: int, edge_factor: int = 16, seed: int = 1, symmetric: bool = False,
               segments: int = 0, mem_bytes: int = 1 << 30, abc=RMAT_ABC, threads: int = 0) -> int:
    """Out-of-core R-MAT (cig_rmat_write): the graph rmat_graph builds with the
    same arguments, written to a CIGG file by source segments; returns m."""
    m = C.c_uint64()
    _check(lib().cig_rmat_write(path.encode(), scale, edge_factor, abc[0], abc[1], abc[2], seed,
                                1 if symmetric else 0, segments, mem_bytes, threads, C.byref(m)))
    return m.value


def read_csr_bin(path: str) -> HostGraph:
    n = C.c_uint32()
    off = C.c_void_p()
    dst = C.c_void_p()
    _check(lib().cig_read_csr_bin(path.encode(), C.byref(n), C.byref(off), C.byref(dst)))
    offsets = _take(off, n.value + 1, np.uint64)
    return HostGraph(n.value, offsets, _take(dst, int(offsets[-1]), np.uint32))


def write_updates_bin(s: "UpdateStream", path: str) -> None:
    """Binary update stream ("CIGU", see cig.h)."""
    sizes = np.ascontiguousarray(s.batch_sizes, dtype=np.uint32)
    adds = np.ascontiguousarray(s.adds, dtype=EDGE_DTYPE)
    dels = np.ascontiguousarray(s.dels, dtype=EDGE_DTYPE)
    _check(lib().cig_write_updates_bin(path.encode(), _ptr(adds) if adds.size else None, adds.size,
                                       _ptr(dels) if dels.size else None, dels.size,
                                       _ptr(sizes) if sizes.size else None, sizes.shape[0]))


def read_updates_bin(path: str) -> "UpdateStream":
    adds = C.c_void_p()
    dels = C.c_void_p()
    sizes = C.c_void_p()
    na = C.c_uint64()
    nd = C.c_uint64()
    nb = C.c_uint32()
    _check(lib().cig_read_updates_bin(path.encode(), C.byref(adds), C.byref(na), C.byref(dels), C.byref(nd),
                                      C.byref(sizes), C.byref(nb)))
    return UpdateStream(_take(adds, na.value, EDGE_DTYPE), _take(dels, nd.value, EDGE_DTYPE),
                        _take(sizes, 2 * nb.value, np.uint32).reshape(nb.value, 2))


def read_updates(update_path: str, batch_path: str) -> UpdateStream:
    adds = C.c_void_p()
    dels = C.c_void_p()
    sizes = C.c_void_p()
    na = C.c_uint64()
    nd = C.c_uint64()
    nb = C.c_uint32()
    _check(lib().cig_read_updates(update_path.encode(), batch_path.encode(), C.byref(adds),
                                  C.byref(na), C.byref(dels), C.byref(nd), C.byref(sizes),
                                  C.byref(nb)))
    return UpdateStream(_take(adds, na.value, EDGE_DTYPE), _take(dels, nd.value, EDGE_DTYPE),
                        _take(sizes, 2 * nb.value, np.uint32).reshape(nb.value, 2))


# ---------------------------------------------------------------- engine
class Engine:
    """One engine per algorithm instance (reference: Engine<..., TApp, ...>)."""

    def __init__(self, algo="sssp", weight_mode=WEIGHT_REF128, weight_seed=0, cache_bytes=0,
                 device=0, pr_epsilon=1e-6, pr_max_rounds=100, flags=0, edges_per_lane=0,
                 split_degree=0, max_rounds=0, n_stream=3, segments=1, tile_len=0,
                 round_blocks_per_cu=0, pr_pull_divisor=0, pull_blocks_per_cu=0, kernel_timing_every=0,
                 repair_delta=0, pool_slack=0, adapt_split=0):
        L = lib()
        cfg = Config()
        _check(L.cig_config_default(C.byref(cfg)))
        cfg.algo = ALGOS[algo] if isinstance(algo, str) else int(algo)
        cfg.weight_mode = weight_mode
        cfg.weight_seed = weight_seed
        cfg.cache_bytes = cache_bytes
        cfg.device = device
        cfg.pr_epsilon = pr_epsilon
        cfg.pr_max_rounds = pr_max_rounds
        cfg.flags = flags
        cfg.edges_per_lane = edges_per_lane
        cfg.split_degree = split_degree
        cfg.max_rounds = max_rounds
        cfg.n_stream = n_stream
        cfg.segments = segments
        cfg.tile_len = tile_len
        cfg.round_blocks_per_cu = round_blocks_per_cu
        cfg.pr_pull_divisor = pr_pull_divisor
        cfg.pull_blocks_per_cu = pull_blocks_per_cu
        cfg.kernel_timing_every = kernel_timing_every
        cfg.repair_delta = repair_delta
        cfg.pool_slack = pool_slack
        cfg.adapt_split = adapt_split
        self.config = cfg
        self._e = C.c_void_p()
        st = L.cig_create(C.byref(cfg), C.byref(self._e))
        if st != OK:
            raise CigError(st, L.cig_last_error(None).decode())
        self.algo = cfg.algo
        self.n = 0
        self.nsrc = 1
        self._server = None
        _LIVE.add(self)

    def close(self):
        srv = getattr(self, "_server", None)
        srv = srv() if srv is not None else None
        if srv is not None:
            srv.stop()  # the server thread owns the engine until it stops
        if getattr(self, "_e", None) and self._e.value:
            lib().cig_destroy(self._e)
            self._e = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _call(self, status):
        if status != OK:
            raise CigError(status, lib().cig_last_error(self._e).decode())

    # LoadGraph
    def load_graph(self, g: HostGraph):
        off = np.ascontiguousarray(g.offsets, dtype=np.uint64)
        dst = np.ascontiguousarray(g.dst, dtype=np.uint32)
        self._call(lib().cig_load_csr(self._e, g.n, _ptr(off), _ptr(dst) if dst.size else None))
        self.n = g.n

    # InitGraph
    def init_graph(self, source=0):
        """One source (an int), or a list of 2..64 BFS/SSSP sources kept
        current together; values()/parents() are then (nsrc, n) arrays."""
        src = np.atleast_1d(np.asarray(source, dtype=np.uint32))
        self._call(lib().cig_init(self._e, _ptr(src), len(src)))
        self.nsrc = len(src) if self.algo in (BFS, SSSP) else 1

    # Start / StartPageRank
    def start(self) -> dict:
        st = BatchStats()
        self._call(lib().cig_converge(self._e, C.byref(st)))
        return st.as_dict()

    def _edges(self, a):
        a = np.ascontiguousarray(a, dtype=EDGE_DTYPE)
        return a, (_ptr(a) if a.size else None)

    # del_edge
    def del_edge(self, dels) -> dict:
        d, p = self._edges(dels)
        st = BatchStats()
        self._call(lib().cig_delete_edges(self._e, p, d.size, C.byref(st)))
        return st.as_dict()

    # add_edge
    def add_edge(self, adds) -> dict:
        a, p = self._edges(adds)
        st = BatchStats()
        self._call(lib().cig_insert_edges(self._e, p, a.size, C.byref(st)))
        return st.as_dict()

    # one batch: deletes then adds (UpdatePageRank for PR)
    def apply_batch(self, dels, adds, staged: bool = False) -> dict:
        d, pd = self._edges(dels)
        a, pa = self._edges(adds)
        st = BatchStats()
        self._call(lib().cig_apply_batch(self._e, pd, d.size, pa, a.size,
                                         BATCH_STAGED if staged else 0, C.byref(st)))
        return st.as_dict()

    # get_update_file
    def load_updates(self, stream: UpdateStream):
        a = np.ascontiguousarray(stream.adds, dtype=EDGE_DTYPE)
        d = np.ascontiguousarray(stream.dels, dtype=EDGE_DTYPE)
        s = np.ascontiguousarray(stream.batch_sizes, dtype=np.uint32)
        self._call(lib().cig_load_updates(self._e, _ptr(a) if a.size else None, a.size,
                                          _ptr(d) if d.size else None, d.size, _ptr(s),
                                          stream.nbatches))

    def apply_next_batch(self, staged: bool = False) -> dict:
        st = BatchStats()
        self._call(lib().cig_apply_next_batch(self._e, BATCH_STAGED if staged else 0,
                                              C.byref(st)))
        return st.as_dict()

    def apply_next_batches(self, count: int, each: bool = False, staged: bool = False):
        """The next `count` resident batches as one group (same results as
        applying them one by one, one mutation and one fixed point): one
        stats dict. each=True: applied one at a time in one library call
        (every intermediate state produced): a list of per-batch dicts."""
        flags = BATCH_STAGED if staged else 0
        if each:
            arr = (BatchStats * count)()
            self._call(lib().cig_apply_next_batches(self._e, count, flags | BATCHES_EACH, arr))
            return [x.as_dict() for x in arr]
        st = BatchStats()
        self._call(lib().cig_apply_next_batches(self._e, count, flags, C.byref(st)))
        return st.as_dict()

    def autotune(self, source: int = 0) -> dict:
        """Picks the fastest launch shape for this graph (results unchanged);
        leaves the engine converged from `source`."""
        r = TuneResult()
        self._call(lib().cig_autotune(self._e, source, C.byref(r)))
        return r.as_dict()

    def query(self, algo: str, sources) -> np.ndarray:
        """BFS / SSSP from up to 64 sources on the live graph: array (nsrc, n)."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        out = np.zeros((len(src), self.n), dtype=np.uint32)
        st = BatchStats()
        self._call(lib().cig_query(self._e, ALGOS[algo], _ptr(src), len(src), _ptr(out), out.nbytes,
                                   C.byref(st)))
        self.last_query_stats = st.as_dict()
        return out

    def query_noalloc(self, algo: str, sources, dist_device: int, dist_bytes: int) -> dict:
        """cig_query_noalloc: the distances go to device memory at address
        `dist_device` (nsrc * n uint32, source-major); nothing is allocated."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        st = BatchStats()
        self._call(lib().cig_query_noalloc(self._e, ALGOS[algo], _ptr(src), len(src),
                                           C.c_void_p(dist_device), dist_bytes, C.byref(st)))
        return st.as_dict()

    def batches_left(self) -> int:
        return int(lib().cig_update_batches_left(self._e))

    # compute_hot_vertices + confirm_candidate_batch + evict/compact/LoadCache
    def cache_refresh(self) -> dict:
        st = BatchStats()
        self._call(lib().cig_cache_refresh(self._e, C.byref(st)))
        return st.as_dict()

    def compact(self, tier) -> None:
        """Packs the 'hbm' (0) or 'host' (1) tier's edge pool now."""
        t = {"hbm": 0, "host": 1}.get(tier, tier)
        self._call(lib().cig_compact(self._e, int(t)))

    def info(self) -> dict:
        gi = GraphInfo()
        self._call(lib().cig_graph_info_get(self._e, C.byref(gi)))
        return gi.as_dict()

    # GatherValue / GatherParent / GatherBuffer
    def gather(self, field: int) -> np.ndarray:
        if field == FIELD_TIER:
            out = np.empty(self.n, dtype=np.uint8)
        elif self.algo == PR and field in (FIELD_VALUE, FIELD_RESIDUAL):
            out = np.empty(self.n, dtype=np.float32)
        elif self.nsrc > 1 and field in (FIELD_VALUE, FIELD_PARENT):
            out = np.empty((self.nsrc, self.n), dtype=np.uint32)  # source-major
        else:
            out = np.empty(self.n, dtype=np.uint32)
        self._call(lib().cig_gather(self._e, field, _ptr(out), out.nbytes))
        return out

    def values(self) -> np.ndarray:
        return self.gather(FIELD_VALUE)

    def parents(self) -> np.ndarray:
        return self.gather(FIELD_PARENT)

    def residuals(self) -> np.ndarray:
        return self.gather(FIELD_RESIDUAL)

    def topology(self) -> HostGraph:
        info = self.info()
        off = np.zeros(self.n + 1, dtype=np.uint64)
        cap = max(int(info["edges"]), 1)
        dst = np.zeros(cap, dtype=np.uint32)
        self._call(lib().cig_gather_topology(self._e, _ptr(off), _ptr(dst), cap))
        return HostGraph(self.n, off, dst[: int(off[-1])])

    def incoming(self) -> HostGraph:
        """The device incoming-row index as CSR (row v = sources of v's in-edges)."""
        info = self.info()
        off = np.zeros(self.n + 1, dtype=np.uint64)
        cap = max(int(info["edges"]), 1)
        src = np.zeros(cap, dtype=np.uint32)
        self._call(lib().cig_gather_incoming(self._e, _ptr(off), _ptr(src), cap))
        return HostGraph(self.n, off, src[: int(off[-1])])


class Server:
    """The engine's serving loop (cig_serve_*): submit updates and queries
    from any thread; results come back in submission order semantics."""

    def __init__(self, engine: "Engine", max_group=32, max_query_sources=64, hbm_reserve=1 << 30,
                 refresh_every=None, max_query_wait_ms=None):
        cfg = ServeConfig()
        lib().cig_serve_config_default(C.byref(cfg))
        cfg.max_group = max_group
        cfg.max_query_sources = max_query_sources
        cfg.hbm_reserve = hbm_reserve
        if refresh_every is not None:
            cfg.refresh_every = refresh_every
        if max_query_wait_ms is not None:
            cfg.max_query_wait_ms = max_query_wait_ms
        self._s = C.c_void_p()
        self._eng = engine
        engine._call(lib().cig_serve_start(engine._e, C.byref(cfg), C.byref(self._s)))
        engine._server = weakref.ref(self)
        self._outs = {}
        _LIVE.add(self)

    def submit_update(self, dels, adds) -> int:
        d = np.ascontiguousarray(dels, dtype=EDGE_DTYPE)
        a = np.ascontiguousarray(adds, dtype=EDGE_DTYPE)
        t = C.c_uint64()
        _check(lib().cig_submit_update(self._s, _ptr(d) if d.size else None, d.size,
                                       _ptr(a) if a.size else None, a.size, C.byref(t)))
        return t.value

    def submit_query(self, algo: str, sources, out=None) -> int:
        """`out` (uint32, at least len(sources)*n elements) receives the
        distances; it is allocated here when not given."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        if out is None:
            out = np.zeros((len(src), self._eng.n), dtype=np.uint32)
        elif out.dtype != np.uint32 or not out.flags.c_contiguous or out.size < len(src) * self._eng.n:
            raise ValueError("out must be a contiguous uint32 array of at least len(sources)*n elements")
        t = C.c_uint64()
        _check(lib().cig_submit_query(self._s, ALGOS[algo], _ptr(src), len(src), _ptr(out), out.nbytes,
                                      C.byref(t)))
        self._outs[t.value] = out  # must outlive the request
        return t.value

    def _collect(self, fn, ticket):
        info = RequestInfo()
        st = fn(self._s, ticket, C.byref(info))
        if st == EPENDING:
            return None
        _check(st)
        d = info.as_dict()
        d["result"] = self._outs.pop(ticket, None)
        return d

    def poll(self, ticket: int):
        return self._collect(lib().cig_poll, ticket)

    def wait(self, ticket: int) -> dict:
        return self._collect(lib().cig_wait, ticket)

    def stop(self):
        if getattr(self, "_s", None) and self._s.value:
            s, self._s = self._s, C.c_void_p()
            _check(lib().cig_serve_stop(s))

    def __del__(self):
        self.stop()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()
