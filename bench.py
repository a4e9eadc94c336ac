#!/usr/bin/env python3
"""Headline benchmark: incremental SSSP batch updates on one MI355X.

Default workload (--config 2, the headline): R-MAT scale 22 (4,194,304
vertices, edge factor 16, a/b/c/d = 0.57/0.19/0.19/0.05, self-loops and
duplicates removed, ids not permuted), source 0, weights
1 + splitmix64(seed ^ (u<<32|v)) % 64 (uniform in [1,64]). Each step is one
update batch of 100,000 deletes sampled uniformly from the current edges and
100,000 new R-MAT inserts, applied by the HIP engine with the whole graph
resident in HBM and the update stream already resident in HBM. Distances are
exact (the parity tests check them bit for bit against the CPU oracle).

Prints ONE JSON line: value = GTEPS (edges scanned by the invalidation,
repair and relax kernels / wall time of the timed batches), with p50/p99
batch latency, updates/s, the batched form (groups of batches in one call),
the roofline of the dominant kernel and a CPU baseline (the oracle's
restatement of the reference CPU path: multigraph mutation + full
recompute, single thread, a bounded sample of the same batches).

--config 1 / 3 / 4 run the other incremental workloads at their own sizes
(BFS S16, CC on a symmetrised S24, PageRank S25) with the same report.
--config 5 is the serving workload: R-MAT scale 27 edge factor 32 under a
16 GiB HBM edge budget (the rest of the rows zero-copy in pinned host
memory), Poisson request arrivals, update requests of log-uniform size,
interleaved 64-source BFS/SSSP queries, all through the serving loop.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TIMING_EVERY = 10

# algo, scale, edge factor, symmetric, deletes/inserts per batch (pairs when
# symmetric), timed per-batch steps, batches per group
CONFIGS = {
    1: dict(algo="bfs", scale=16, ef=16, sym=False, batch=1000, steps=100, group=10,
            what="incremental BFS, R-MAT scale 16 edge factor 16, source 0, batches of 1000 deletes + "
                 "1000 inserts"),
    2: dict(algo="sssp", scale=22, ef=16, sym=False, batch=100_000, steps=100, group=10,
            what="incremental SSSP, R-MAT scale 22 edge factor 16 (a,b,c,d=0.57,0.19,0.19,0.05, deduped, "
                 "unpermuted), source 0, uint32 weights uniform in [1,64], batches of 100000 deletes + "
                 "100000 inserts, graph and stream resident in HBM"),
    3: dict(algo="cc", scale=24, ef=8, sym=True, batch=1_000_000, steps=10, group=10, autotune=True,
            what="incremental min-label CC, symmetrised R-MAT scale 24 (8 undirected samples per vertex, "
                 "both directions stored: ~268M directed edges), batches of 1M symmetric delete pairs + 1M "
                 "symmetric insert pairs, autotuned launch shape"),
    4: dict(algo="pr", scale=25, ef=16, sym=False, batch=500_000, steps=10, group=10,
            what="incremental PageRank (alpha 0.85, epsilon 1e-7), R-MAT scale 25 edge factor 16, "
                 "batches of 500000 deletes + 500000 inserts"),
    5: dict(algo="sssp", scale=27, ef=32, sym=False, batch=0, steps=0, group=32, serve=True,
            what="serving: incremental SSSP on R-MAT scale 27 edge factor 32 with a 16 GiB HBM edge budget "
                 "(other rows zero-copy in pinned host memory), Poisson arrivals, update requests of "
                 "log-uniform size in [1k, 1M] (half deletes, half inserts), interleaved 64-source BFS/SSSP "
                 "queries, continuous batching, admission by free HBM"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info():
    """CPU model string, hardware threads and the threads this process may use."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu": model, "hardware_concurrency": os.cpu_count(), "nproc": usable}


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


def stage_bytes(st, stage, pr=False):
    """Algorithmic bytes of one expansion pass. Traversal: 8 B per scanned
    edge (4-B destination id + 4-B destination state touch) + 28 B per
    expanded row (16-B descriptor, 8-B state word, 4-B worklist entry).
    PageRank scatter (stage 2): 8 B per edge (id + residual atomic) + 40 B per
    consumed vertex. PageRank pull rounds (stage 1): 8 B per incoming edge
    (4-B source id + 4-B contribution read) + 24 B per gathering vertex
    (16-B incoming-row descriptor, residual read and write). Weights are not
    stored."""
    per_v = (24.0 if stage == 1 else 40.0) if pr else 28.0
    return 8.0 * st["stage_edges"][stage] + per_v * st["stage_vertices"][stage]


def read_traffic(path):
    """Measured HBM bytes per launch of the dominant kernel, from the
    committed rocprofv3 --pmc summary (line 'traffic_bytes_per_launch = X')."""
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1].split()[0])
    except OSError:
        pass
    return None


def cpu_baseline(g, stream, batches, algo, wseed, budget_s):
    """Reference CPU path on the same batches, single-threaded: chunk-store
    mutation + full recompute (label-correcting SSSP as SSSPHostNaive with
    the engine's +1, queue BFS, CCCheck flood, FP64 Jacobi PageRank), the
    reference's --check oracles restated in oracle/."""
    from oracle import oracle as O
    og = O.Graph(g.n, g.offsets, g.dst)
    done = 0
    scanned = 0
    t_total = 0.0
    for b in range(batches):
        dels, adds = stream.batch(b)
        t0 = time.perf_counter()
        og.apply_deletes(dels["u"], dels["v"])
        og.apply_adds(adds["u"], adds["v"])
        if algo == "sssp":
            _, sc = og.sssp_label_correcting(0, O.W_HASH64, wseed, with_count=True)
        elif algo == "bfs":
            og.bfs(0)
            sc = og.edge_count()
        elif algo == "cc":
            og.cc()
            sc = og.edge_count()
        else:  # three sweeps of the FP64 Jacobi recompute (a full one needs ~170)
            og.pagerank(max_rounds=3)
            sc = og.edge_count() * 3
        t_total += time.perf_counter() - t0
        scanned += sc
        done += 1
        log(f"cpu baseline batch {b}: {t_total / done * 1e3:.0f} ms/batch")
        if t_total >= budget_s:
            break
    what = {"sssp": "label-correcting SSSP", "bfs": "queue BFS", "cc": "flood-fill CC",
            "pr": "3 sweeps of FP64 Jacobi PageRank"}[algo]
    return {
        "value": round(scanned / t_total / 1e9, 4),
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": (f"{done} batches of the same stream: mutation + {what} "
                   f"recompute, {t_total / done * 1e3:.1f} ms/batch, "
                   f"{scanned / done / 1e6:.1f} M edges scanned/batch"),
        "ms_per_batch": round(t_total / done * 1e3, 3),
    }


def cpu_mutation_pool(g, stream, batches, budget_s):
    """The reference's CPU batch maintenance as its worker pool runs it:
    per-source chunk mutation with min(20, hardware threads) workers
    (source_local_chunk_store.h:116-121, 218-391), restated in the oracle,
    on the same batches, timed alone (no recompute)."""
    from oracle import oracle as O
    workers = min(20, os.cpu_count() or 1)
    og = O.Graph(g.n, g.offsets, g.dst)
    t_total, done, recs = 0.0, 0, 0
    for b in range(batches):
        dels, adds = stream.batch(b)
        t0 = time.perf_counter()
        og.apply_batch_mt(dels["u"], dels["v"], adds["u"], adds["v"], workers)
        t_total += time.perf_counter() - t0
        done += 1
        recs += len(dels) + len(adds)
        if t_total >= budget_s:
            break
    return {"workers": workers, "batches": done, "ms_per_batch": round(t_total / done * 1e3, 3),
            "updates_per_s": round(recs / t_total, 1)}


def multi_source_form(g, stream, batches, algo, wseed, single_p50):
    """One engine keeping 64 sources current through every batch (cig_init
    with 64 sources): per-batch latency and source-updates/s against 64
    single-source engines at the single-source p50 (every batch's state is
    produced for every source in both forms)."""
    import coincgraph_amd as cg
    k = 64
    sources = (np.arange(k, dtype=np.uint64) * 65537 % g.n).astype(np.uint32)
    sources[0] = 0
    kw = dict(weight_mode=cg.WEIGHT_HASH64, weight_seed=wseed) if algo == "sssp" else {}
    with cg.Engine(algo, **kw) as e:
        e.load_graph(g)
        e.init_graph(sources)
        st0 = e.start()
        e.load_updates(stream)
        lat = []
        for _ in range(batches):
            lat.append(e.apply_next_batch()["total_ms"])
    p50 = percentile(lat, 50)
    per_batch_updates = 2 * int(stream.batch_sizes[0][0])
    return {"sources": k, "batches": batches, "initial_ms": round(st0["total_ms"], 3),
            "latency_ms": {"p50": round(p50, 4), "p99": round(percentile(lat, 99), 4), "samples": len(lat)},
            "source_updates_per_s": round(k * per_batch_updates / (p50 * 1e-3), 1),
            "speedup_vs_single_source_engines": round(k * single_p50 / p50, 2)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=sorted(CONFIGS))
    ap.add_argument("--steps", type=int, default=None, help="timed per-batch steps")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=int, default=None)
    ap.add_argument("--edge-factor", type=int, default=None)
    ap.add_argument("--batch", type=int, default=None, help="deletes and inserts per batch")
    ap.add_argument("--seed", type=int, default=2022)
    ap.add_argument("--weight-seed", type=int, default=64)
    ap.add_argument("--staged", action="store_true", help="reference two-stage batches")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true", help="verify the final values on the CPU")
    ap.add_argument("--group", type=int, default=None,
                    help="batched form: batches per group (cig_apply_next_batches)")
    ap.add_argument("--group-steps", type=int, default=None, help="batched form: timed groups (0 = skip)")
    ap.add_argument("--autotune", action="store_true", help="pick the launch shape first (cig_autotune)")
    ap.add_argument("--no-autotune", action="store_true", help="skip the autotuner (config 3 runs it by default)")
    ap.add_argument("--no-multi-source", action="store_true", help="skip the 64-source engine (configs 1, 2)")
    ap.add_argument("--ordered", action="store_true",
                    help="SSSP: delta-ordered rounds (CIG_CFG_ORDERED_REPAIR, the reference's ordered executor)")
    ap.add_argument("--cc-union", action="store_true",
                    help="CC: union-find repair (CIG_CFG_CC_UNION_REPAIR, the reference's union repair)")
    ap.add_argument("--graph", action="store_true",
                    help="replay each batch's fixed-length parts from captured HIP graphs (CIG_CFG_GRAPH_BATCHES)")
    ap.add_argument("--adapt-split", type=int, default=0,
                    help="frontier rounds keep rows of up to this many slots whole when waves take few "
                         "items (cig_config.adapt_split; 0 = the library default, 1 = off)")
    ap.add_argument("--n-stream", type=int, default=0,
                    help="compute streams (cig_config.n_stream: 1 one stream, 2 grouping beside the "
                         "invalidation waves, 3 grouping and delete claims beside them; 0 = default 3)")
    ap.add_argument("--round-blocks", type=int, default=0,
                    help="workgroups per CU of the round kernels (cig_config.round_blocks_per_cu; 0 = default)")
    ap.add_argument("--split", type=int, default=0, help="tile split degree (cig_config.split_degree; 0 = default)")
    ap.add_argument("--tile-len", type=int, default=0, help="slots per tile (cig_config.tile_len; 0 = default)")
    ap.add_argument("--pull-divisor", type=int, default=0,
                    help="PageRank: rounds with at least n/this many frontier vertices gather "
                         "(cig_config.pr_pull_divisor; 0 = default)")
    ap.add_argument("--pull-blocks", type=int, default=0,
                    help="PageRank: workgroups per CU of the pull kernels (0 = default)")
    ap.add_argument("--edges-per-lane", type=int, default=0,
                    help="edges each lane keeps in flight: 4, 8 or 16 (0 = default)")
    ap.add_argument("--pr-push", action="store_true",
                    help="PageRank: scatter every round with atomics (keeps no incoming rows)")
    ap.add_argument("--serve-seconds", type=float, default=30.0, help="serving: arrival window (s)")
    ap.add_argument("--rate", type=float, default=20.0, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--query-every", type=int, default=200,
                    help="serving: every k-th request is a 64-source query (0 = none)")
    ap.add_argument("--budget-gb", type=float, default=16.0, help="serving: HBM edge budget (GiB)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the engine has no multi-GPU path "
              f"(--gpus {args.gpus} requested). Run with --gpus 1.", file=sys.stderr)
        return 2

    c = CONFIGS[args.config]
    if c.get("serve"):
        return serve_bench(args, c)
    algo = c["algo"]
    scale = args.scale or c["scale"]
    ef = args.edge_factor or c["ef"]
    batch = args.batch or c["batch"]
    steps = args.steps if args.steps is not None else c["steps"]
    group = args.group or c["group"]
    group_steps = args.group_steps if args.group_steps is not None else (10 if args.config <= 2 else 1)
    # at least 100 latency samples where batches are cheap (configs 1 and 2),
    # taken after the timed steps when --steps asks for fewer
    lat_extra = max(0, 100 - steps) if args.config in (1, 2) else 0
    ms_batches = 20 if args.config in (1, 2) and not args.no_multi_source else 0
    pr = algo == "pr"
    mult = 2 if c["sym"] else 1  # directed records per requested update

    import coincgraph_amd as cg

    t0 = time.perf_counter()
    g = cg.rmat_graph(scale, ef, seed=args.seed, symmetric=c["sym"])
    log(f"graph: n={g.n} m={g.m} ({time.perf_counter() - t0:.1f} s)")
    nb = args.warmup + steps + lat_extra + group * group_steps
    t0 = time.perf_counter()
    stream = cg.rmat_stream(g, scale, nb, batch, batch, seed=args.seed + 1, symmetric=c["sym"])
    log(f"stream: {nb} batches of {batch}+{batch}{' pairs' if c['sym'] else ''} "
        f"({time.perf_counter() - t0:.1f} s)")

    # round kernels are timed (an event pair around each launch) in one batch
    # of every TIMING_EVERY: the pairs idle the stream ~10 us each, so timing
    # every batch would inflate the latency it reports
    # (PageRank's round kernels run for milliseconds: every batch is timed)
    kw = dict(weight_mode=cg.WEIGHT_HASH64, weight_seed=args.weight_seed, flags=cg.CFG_KERNEL_TIMING,
              kernel_timing_every=1 if pr else TIMING_EVERY)
    if c["sym"]:
        kw["flags"] |= cg.CFG_SYMMETRIC  # undirected graph: outgoing rows double as incoming rows
    if args.ordered:
        kw["flags"] |= cg.CFG_ORDERED_REPAIR
    if args.cc_union:
        kw["flags"] |= cg.CFG_CC_UNION_REPAIR
    if args.graph:
        kw["flags"] |= cg.CFG_GRAPH_BATCHES
    if args.adapt_split:
        kw["adapt_split"] = args.adapt_split
    if args.round_blocks:
        kw["round_blocks_per_cu"] = args.round_blocks
    if args.n_stream:
        kw["n_stream"] = args.n_stream
    if args.split:
        kw["split_degree"] = args.split
    if args.tile_len:
        kw["tile_len"] = args.tile_len
    if args.pull_divisor:
        kw["pr_pull_divisor"] = args.pull_divisor
    if args.pull_blocks:
        kw["pull_blocks_per_cu"] = args.pull_blocks
    if args.edges_per_lane:
        kw["edges_per_lane"] = args.edges_per_lane
    if pr:
        kw.update(pr_epsilon=1e-7, pr_max_rounds=1000)
        if args.pr_push:
            kw["flags"] |= cg.CFG_PR_PUSH_ONLY
    eng = cg.Engine(algo, **kw)
    eng.load_graph(g)
    tune = None
    if (args.autotune or c.get("autotune")) and not args.no_autotune and not pr:
        tune = eng.autotune(0)
        log(f"autotune: {tune}")
    else:
        eng.init_graph(0)
        st0 = eng.start()
        log(f"initial fixed point: {st0['total_ms']:.2f} ms, rounds={st0['rounds']}, "
            f"E_scan={st0['edges_scanned']}")
    eng.load_updates(stream)
    for _ in range(args.warmup):
        eng.apply_next_batch(staged=args.staged)
    log(f"warmup: {args.warmup} batches")

    # The timed batches go through one library call that applies them one at
    # a time (CIG_BATCHES_EACH): each batch runs alone, with its own mutation
    # and fixed point, and every intermediate state is produced; only the
    # Python interpreter between batches is left out.
    t0 = time.perf_counter()
    stats = eng.apply_next_batches(steps, each=True, staged=args.staged)
    wall = time.perf_counter() - t0
    log(f"timed batches done: {steps}")
    info = eng.info()
    lat_stats = list(stats)
    for _ in range(lat_extra):  # latency samples only: not in the timed value
        lat_stats.append(eng.apply_next_batch(staged=args.staged))

    # Batched form: the next batches applied `group` at a time (one mutation
    # and one fixed point per group; same final state as one by one).
    grouped = None
    if group_steps:
        gstats = []
        t1 = time.perf_counter()
        for _ in range(group_steps):
            gstats.append(eng.apply_next_batches(group))
        gwall = time.perf_counter() - t1
        glat = [x["total_ms"] for x in gstats]
        gup = group_steps * group * 2 * batch * mult
        grouped = {
            "form": "coalesced groups: one mutation and one fixed point per group; the states between "
                    "the batches of a group are not produced",
            "batches_per_group": group,
            "groups": group_steps,
            "updates_per_s": round(gup / gwall, 1),
            "ms_per_group": {"p50": round(percentile(glat, 50), 4), "p99": round(percentile(glat, 99), 4)},
            "speedup_vs_per_batch": round((gup / gwall) / (steps * 2 * batch * mult / wall), 3),
        }
        log(f"batched form: {grouped}")

    lat = [s["total_ms"] for s in lat_stats]
    e_scan = sum(s["edges_scanned"] for s in stats)
    updates = steps * 2 * batch * mult
    gteps = e_scan / wall / 1e9

    # stage markers are recorded on the sampled batches only (kernel_timing_every)
    staged = [s for s in stats if s["round_launches"]] or stats
    stage_ms = {"propagate": sum(s["propagate_ms"] for s in staged) / len(staged),
                "repair": sum(s["repair_ms"] for s in staged) / len(staged),
                "mutation": sum(s["mutate_ms"] for s in staged) / len(staged),
                "invalidation": sum(s["invalidate_ms"] for s in staged) / len(staged)}
    # Dominant kernel: the propagation round kernel (k_relax_round; for
    # PageRank k_pr_pull, the gather form of dense rounds, or k_pr_scatter
    # when no round gathered). Its duration comes from the HIP event pair the
    # engine records around every launch on its own stream; its algorithmic
    # bytes are its pass's (stage_bytes) spread over all its launches (no-op
    # lookahead rounds, and for k_pr_pull the sparse rounds it leaves to the
    # scatter, included, as rocprofv3 counts them).
    # Only the batches whose round kernels were timed enter (one in TIMING_EVERY;
    # every batch for PageRank).
    timed = [s for s in stats if s["round_launches"]]
    launches = sum(s["round_launches"] for s in timed)
    kern_ms = sum(s["round_kernel_ms"] for s in timed)
    pull = pr and sum(s["pull_rounds"] for s in stats) > 0
    prop_bytes = sum(stage_bytes(s, 1 if pull else 2, pr) for s in timed)
    bytes_per_launch = prop_bytes / max(launches, 1)
    ms_per_launch = kern_ms / max(launches, 1)
    achieved = bytes_per_launch / (ms_per_launch * 1e-3) / 1e9 if ms_per_launch > 0 else 0.0
    kernel = ("k_pr_pull" if pull else "k_pr_scatter") if pr else "k_relax_round"
    traffic = read_traffic(os.path.join(ROOT, "profiles", f"traffic_{kernel}_config{args.config}.txt"))

    if args.check:
        from oracle import oracle as O
        og = O.Graph(g.n, g.offsets, g.dst)
        for b in range(nb):
            d, a = stream.batch(b)
            og.apply_deletes(d["u"], d["v"])
            og.apply_adds(a["u"], a["v"])
        if pr:
            ref = og.pagerank()[0]
            l1n = float(np.abs(eng.values().astype(np.float64) - ref).sum() / np.abs(ref).sum())
            ok = l1n <= 1e-6
            log(f"check vs FP64 oracle after {nb} batches: normalised L1 {l1n:.3g} ({'ok' if ok else 'FAIL'})")
        else:
            exp = {"sssp": lambda: og.sssp(0, O.W_HASH64, args.weight_seed), "bfs": lambda: og.bfs(0),
                   "cc": og.cc}[algo]()
            ok = np.array_equal(eng.values(), exp)
            log(f"check vs oracle after {nb} batches: {'ok' if ok else 'MISMATCH'}")
        if not ok:
            return 1

    eng.close()
    multi = None
    if ms_batches and algo in ("sssp", "bfs"):
        multi = multi_source_form(g, stream, ms_batches, algo, args.weight_seed, percentile(lat, 50))
        log(f"multi-source form: {multi}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(g, stream, nb, algo, args.weight_seed, args.cpu_seconds)
        cpu["mutation_pool"] = cpu_mutation_pool(g, stream, nb, args.cpu_seconds / 4)
        log(f"cpu mutation pool: {cpu['mutation_pool']}")

    default_shape = (args.scale, args.edge_factor, args.batch) == (None, None, None)
    out = {
        "metric": "batch-update p50/p99 latency + GTEPS (% HBM peak), incremental SSSP, 1 GPU",
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if pr else "u32",
        "data": "synthetic",
        "config": {
            "workload": c["what"] if default_shape else
            f"{algo} R-MAT scale {scale} edge factor {ef}, batches of {batch} deletes + {batch} inserts",
            "config": args.config, "vertices": g.n, "edges": g.m, "batch_deletes": batch,
            "batch_inserts": batch, "symmetric_pairs": c["sym"], "mode": "staged" if args.staged else "fused",
            "executor": ("delta-ordered" if args.ordered else "union-find" if args.cc_union else "default")
            + (", graph-replayed batches" if args.graph else ""),
        },
        "latency_ms": {"p50": round(percentile(lat, 50), 4), "p99": round(percentile(lat, 99), 4),
                       "mean": round(float(np.mean(lat)), 4), "max": round(max(lat), 4), "samples": len(lat)},
        "updates_per_s": round(updates / wall, 1),
        "edges_scanned_per_batch": int(e_scan / steps),
        "stage_ms_per_batch": {k: round(v, 4) for k, v in stage_ms.items()},
        "repair_kernel_ms_per_batch": round(sum(s["scan_kernel_ms"] for s in staged) / len(staged), 4),
        "affected_per_batch": round(float(np.mean([s["affected"] for s in stats])), 1),
        "rounds_per_batch": round(float(np.mean([s["rounds"] for s in stats])), 1),
        **({"invalidation_rounds_per_batch": round(float(np.mean([s["inval_rounds"] for s in stats])), 1)}
           if not pr else {}),
        **({"pull_rounds_per_batch": round(float(np.mean([s["pull_rounds"] for s in stats])), 1)} if pr else {}),
        "batched_form": grouped,
        "multi_source_form": multi,
        "roofline": {
            "kernel": kernel,
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": round(bytes_per_launch, 1),
            "ms_per_launch": round(ms_per_launch, 5),
            "launches_per_batch": round(launches / max(len(timed), 1), 2),
            "timed_batches": len(timed),
        },
        "cpu_baseline": cpu,
        "autotune": tune,
        "host": host_info(),
        "engine": {k: info[k] for k in ("block_size", "edges_per_lane", "split_degree", "cu_count",
                                        "hbm_pool_slots")},
    }
    print(json.dumps(out), flush=True)
    return 0


def slowest_groups(res, k=6):
    """The k longest services (one entry per group): when the group started
    (s after the first submission), its service and queueing time, its size."""
    t0 = min(r["submit_ms"] for _, _, r in res)
    seen, out = set(), []
    for kind, _, r in res:
        key = (r["start_ms"], r["done_ms"])
        if key in seen:
            continue
        seen.add(key)
        st = r["stats"]
        out.append({"kind": kind, "start_s": round((r["start_ms"] - t0) / 1e3, 3),
                    "service_ms": round(r["done_ms"] - r["start_ms"], 1),
                    "waited_ms": round(r["start_ms"] - r["submit_ms"], 1), "group_size": r["group_size"],
                    "device_ms": round(st["total_ms"], 1), "rounds": st["rounds"],
                    "records": st["deletes_applied"] + st["missing_deletes"] + st["adds_applied"],
                    "relocated_rows": st["relocated_rows"], "compactions_so_far": st["pool_compactions"]})
    out.sort(key=lambda x: -x["service_ms"])
    return out[:k]


def serve_bench(args, c):
    """Open-loop serving run: requests arrive by a Poisson process and go
    through the engine's serving loop (cig_serve_*), which coalesces queued
    updates into groups and admits queries by free HBM. Latency = done -
    submit as the server clocks it; updates/s = update records applied over
    the span from the first submission to the last completion."""
    import coincgraph_amd as cg

    scale = args.scale or c["scale"]
    ef = args.edge_factor or c["ef"]
    rng = np.random.default_rng(args.seed + 5)
    plan, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / args.rate))
        if t > args.serve_seconds:
            break
        i = len(plan)
        if args.query_every and i % args.query_every == args.query_every - 1:
            plan.append(("query", t, "bfs" if (i // args.query_every) % 2 == 0 else "sssp"))
        else:
            k = int(round(math.exp(rng.uniform(math.log(1e3), math.log(1e6)))))
            plan.append(("update", t, k // 2, k - k // 2))
    nd = sum(p[2] for p in plan if p[0] == "update")
    na = sum(p[3] for p in plan if p[0] == "update")
    nq = sum(1 for p in plan if p[0] == "query")
    log(f"plan: {len(plan)} requests over {args.serve_seconds:.0f} s ({nq} queries), {nd}+{na} update records")

    t0 = time.perf_counter()
    g = cg.rmat_graph(scale, ef, seed=args.seed)
    log(f"graph: n={g.n} m={g.m} ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    stream = cg.rmat_stream(g, scale, 1, nd, na, seed=args.seed + 1)
    dels_all, adds_all = stream.batch(0)
    log(f"update records ({time.perf_counter() - t0:.1f} s)")
    budget = int(args.budget_gb * 2**30)
    eng = cg.Engine("sssp", weight_mode=cg.WEIGHT_HASH64, weight_seed=args.weight_seed, cache_bytes=budget)
    t0 = time.perf_counter()
    eng.load_graph(g)
    n, m0 = g.n, g.m
    del g
    info = eng.info()
    log(f"load ({time.perf_counter() - t0:.1f} s): {info}")
    eng.init_graph(0)
    st = eng.start()
    log(f"initial fixed point: {st['total_ms']:.1f} ms, rounds={st['rounds']}")
    probe_bytes = 1 << 30
    probe_ms = cg.host_read_probe(probe_bytes, iters=5)
    probe_gbs = probe_bytes / (probe_ms * 1e-3) / 1e9
    log(f"host-link probe: {probe_gbs:.1f} GB/s zero-copy read")
    # shared by every query (contents unused); touched once here so that the
    # first query's copy does not pay the page faults of a fresh 64*n*4-B buffer
    qout = np.empty(64 * n, dtype=np.uint32)
    qout.fill(0)
    sources = (np.arange(64, dtype=np.uint64) * 2097143 % n).astype(np.uint32)

    srv = cg.Server(eng, max_group=c["group"], max_query_sources=64, hbm_reserve=1 << 30)
    tickets = []
    d0 = a0 = 0
    start = time.perf_counter()
    for i, p in enumerate(plan):
        wait = p[1] - (time.perf_counter() - start)
        if wait > 0:
            time.sleep(wait)
        if p[0] == "update":
            tk = srv.submit_update(dels_all[d0:d0 + p[2]], adds_all[a0:a0 + p[3]])
            d0 += p[2]
            a0 += p[3]
            tickets.append(("update", tk, p[2] + p[3]))
        else:
            tk = srv.submit_query(p[2], sources, out=qout)
            tickets.append((p[2], tk, 64))
        if (i + 1) % 100 == 0:
            log(f"submitted {i + 1}/{len(plan)}")
    done = {}
    last = time.perf_counter()
    while len(done) < len(tickets):
        for kind, tk, size in tickets:
            if tk not in done:
                r = srv.poll(tk)
                if r is not None:
                    r.pop("result", None)
                    done[tk] = (kind, size, r)
        if time.perf_counter() - last > 10:
            last = time.perf_counter()
            log(f"completed {len(done)}/{len(tickets)}")
            if time.perf_counter() - start > args.serve_seconds + 300:
                log("requests still pending 300 s after the last arrival: giving up")
                return 1
        time.sleep(0.01)
    srv.stop()
    # cache admission on the activity the serving run left (age, histogram,
    # threshold, evict/load on the device), then once more with no new activity
    refresh = []
    for _ in range(2):
        rs = eng.cache_refresh()
        refresh.append({"ms": round(rs["total_ms"], 2), "rows_moved": rs["relocated_rows"]})
    log(f"cache refresh: {refresh}")
    end_info = eng.info()
    eng.close()

    res = list(done.values())
    for kind, size, r in res:
        if r["status"] != 0:
            log(f"request failed: {kind} status {r['status']}")
            return 1
    upd = [r for kind, _, r in res if kind == "update"]
    lat_u = [r["done_ms"] - r["submit_ms"] for r in upd]
    span_ms = max(r["done_ms"] for _, _, r in res) - min(r["submit_ms"] for _, _, r in res)
    recs = sum(size for kind, size, _ in res if kind == "update")
    # every request of a coalesced group carries the group's statistics
    host_edges = sum(r["stats"]["host_edges_scanned"] / max(1, r["group_size"]) for _, _, r in res)
    dev_ms = sum(r["stats"]["total_ms"] / max(1, r["group_size"]) for _, _, r in res)
    queries = {}
    for algo in ("bfs", "sssp"):
        lq = [r["done_ms"] - r["submit_ms"] for kind, _, r in res if kind == algo]
        if lq:
            queries[algo] = {"count": len(lq), "p50_ms": round(percentile(lq, 50), 1),
                             "p99_ms": round(percentile(lq, 99), 1),
                             "service_ms_mean": round(float(np.mean([r["done_ms"] - r["start_ms"] for kind, _, r in res
                                                                      if kind == algo])), 1),
                             "device_ms_mean": round(float(np.mean([r["stats"]["total_ms"] for kind, _, r in res
                                                                     if kind == algo])), 1)}
    groups = [r["group_size"] for r in upd]
    cold = 1.0 - end_info["cached_edges"] / max(1, end_info["edges"])
    out = {
        "metric": "serving p50/p99 update latency + updates/s, 1 GPU",
        "value": round(recs / (span_ms * 1e-3), 1),
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": len(plan),
        "warmup": 0,
        "ms_per_step": round(span_ms / max(1, len(plan)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": c["what"], "config": 5, "vertices": n, "edges": m0,
                   "budget_bytes": budget, "rate_per_s": args.rate, "window_s": args.serve_seconds,
                   "query_every": args.query_every, "query_sources": 64},
        "latency_ms": {"p50": round(percentile(lat_u, 50), 3), "p99": round(percentile(lat_u, 99), 3),
                       "mean": round(float(np.mean(lat_u)), 3), "max": round(float(np.max(lat_u)), 3)},
        "updates": {"requests": len(upd), "records": recs, "mean_group": round(float(np.mean(groups)), 2),
                    "service_ms": {"p50": round(percentile([r["done_ms"] - r["start_ms"] for r in upd], 50), 3),
                                   "p99": round(percentile([r["done_ms"] - r["start_ms"] for r in upd], 99), 3)}},
        "queries": queries,
        "slowest_groups": slowest_groups(res),
        "cache_refresh": refresh,
        "host_link": {"cold_edge_fraction": round(cold, 4),
                      "probe_gbs": round(probe_gbs, 1),
                      "request_gbs": round(host_edges * 4 / (dev_ms * 1e-3) / 1e9, 2) if dev_ms else None,
                      "host_edges_scanned": int(host_edges)},
        "roofline": None,
        "cpu_baseline": None,
        "host": host_info(),
        "engine": {k: info[k] for k in ("hbm_pool_slots", "host_pool_slots", "cached_rows", "cached_edges",
                                        "device_bytes", "host_pinned_bytes")},
        "compactions": {"total": end_info["compactions"], "in_place": end_info["inplace_compactions"],
                        "host_tier": end_info["host_compactions"], "ms_hbm": end_info["compaction_ms_hbm"],
                        "ms_host": end_info["compaction_ms_host"],
                        "refreshes_deferred": end_info["refreshes_deferred"]},
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
