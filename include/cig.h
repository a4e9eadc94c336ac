/*
 * cig.h — C-ABI of the MI355X incremental graph engine (libcig.so).
 *
 * The reference has no ABI: its engine is a header-only C++ template
 * (sepgraph::engine::Engine, include/framework/framework.cuh:456-3505)
 * instantiated per sample with an AppBase policy
 * (include/framework/variants/api.cuh:22-174). Each entry point below names
 * the Engine method (file:line) whose behaviour it replaces. The app policy
 * becomes the `algo` field: the four shipped policies (BFS, SSSP, CC, PR) are
 * compiled into the library as HIP kernels.
 *
 * Conventions
 *   - Every function returns a cig_status (0 = CIG_OK). No exception or abort
 *     crosses the ABI; cig_last_error() describes the last failure.
 *   - The engine owns all device and pinned host memory. Caller buffers are
 *     borrowed for the duration of the call only (host pointers unless the
 *     name says _device).
 *   - One engine is driven by one host thread at a time (the reference's
 *     engine is not thread-safe either, framework.cuh; see cig_serve_* for
 *     the concurrent request loop).
 *   - Vertex ids are dense uint32 in [0, n). Unreachable / unset = 0xFFFFFFFF.
 */
#ifndef CIG_H
#define CIG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CIG_ABI_VERSION 1u
#define CIG_INF 0xFFFFFFFFu
#define CIG_MAX_SOURCES 64u

typedef enum cig_status {
    CIG_OK = 0,
    CIG_EINVAL = 1,   /* bad argument (reference: std::invalid_argument) */
    CIG_ENOMEM = 2,   /* device or pinned host allocation failed */
    CIG_EDEVICE = 3,  /* HIP runtime error (reference: GROUTE_CUDA_CHECK abort) */
    CIG_ESTATE = 4,   /* call out of order (reference: protocol abort) */
    CIG_EIO = 5,      /* file read/parse error */
    CIG_ELIMIT = 6,   /* a documented size limit was exceeded */
    CIG_EINTERNAL = 7, /* internal invariant violated (reference: std::abort) */
    CIG_EPENDING = 8   /* serving: the request has not finished yet */
} cig_status;

/* App policy (reference: the AppBase subclass a sample instantiates). */
typedef enum cig_algo {
    CIG_ALGO_BFS = 0,  /* samples/hybrid_bfs/hybrid_bfs.cu:36-192, weight 1 */
    CIG_ALGO_SSSP = 1, /* samples/hybrid_sssp/hybrid_sssp.cu:36-187 */
    CIG_ALGO_CC = 2,   /* samples/hybrid_cc/hybrid_cc.cu:36-186, directed min label */
    CIG_ALGO_PR = 3    /* samples/hybrid_pr/hybrid_pr.cu:33-123, signed residual push */
} cig_algo;

/* SSSP edge weight. Weights are never stored: they are a function of (u,v)
 * as in the reference (api.cuh:26-28; framework.cuh:1583 overwrites the
 * update file's weights). */
typedef enum cig_weight_mode {
    CIG_WEIGHT_REF128 = 0, /* (u+v)%128+1, the reference's TraversalEdgeWeight */
    CIG_WEIGHT_HASH64 = 1  /* 1 + splitmix64(seed ^ (u<<32|v)) % 64, uniform in [1,64] */
} cig_weight_mode;

/* Update record: the reference's WEdgePair<u32,u32,u32> (edgepair.h:17-30).
 * w is accepted for layout compatibility and ignored. */
typedef struct cig_edge {
    uint32_t u, v, w;
} cig_edge;

enum {
    CIG_CFG_COOPERATIVE = 1u << 0, /* run fixed points as one persistent
                                      cooperative launch with grid barriers
                                      instead of chained round kernels */
    CIG_CFG_NO_HOTNESS = 1u << 1,  /* do not count row activity (never counted
                                      without a cache budget, cache_bytes = 0) */
    CIG_CFG_SCAN_REPAIR = 1u << 2, /* keep no incoming-row index: BFS/SSSP/CC repair
                                      deletions by scanning every row, PR scatters
                                      every round (half the edge memory) */
    CIG_CFG_KERNEL_TIMING = 1u << 3, /* bracket every propagation round kernel with
                                        HIP events (cig_batch_stats.round_kernel_ms) */
    CIG_CFG_PR_PUSH_ONLY = 1u << 4, /* PR: keep no incoming-row index; every round
                                      scatters with atomics (the reference's form) */
    CIG_CFG_SYMMETRIC = 1u << 5,   /* the caller asserts an undirected graph: every edge (u,v)
                                      has its reverse (v,u) as often, and every batch keeps it so
                                      (the samples' CC input). The outgoing rows then serve as
                                      the incoming rows: no mirrored updates, half the mutation
                                      work and edge memory. cig_load_csr rejects a graph whose
                                      in- and out-degrees differ. */
    CIG_CFG_ORDERED_REPAIR = 1u << 6, /* SSSP: delta-ordered rounds (the reference's
                                      CG_ORDERED_REPAIR executor, ordered_gpu_repair.cuh:117-160):
                                      each round expands only the frontier rows whose value /
                                      repair_delta is the round's smallest, the rest wait. Fewer
                                      redundant relaxations on large-diameter graphs; same results */
    CIG_CFG_GRAPH_BATCHES = 1u << 8, /* BFS/SSSP/CC single-source batches: the fixed-length parts of a
                                      batch (the mutation with the invalidation seeds; the reset,
                                      pull repair and insertion seeds) are replayed from two captured
                                      HIP graphs, the kernels reading the batch's records through a
                                      device reference; the round chains stay host-driven. Same
                                      results; ignored for PR, several sources, staged or grouped
                                      batches and CIG_CFG_SCAN_REPAIR */
    CIG_CFG_CC_UNION_REPAIR = 1u << 7, /* CC on an undirected graph (needs CIG_CFG_SYMMETRIC): the
                                      reference's union repair (cc_union_repair.cuh:9-137) in place
                                      of label propagation. A delete re-derives the components it
                                      touches by sampled union-find (the largest sampled component
                                      skips hooking), an insert unites two components; labels are
                                      the same, parents are not kept (CIG_INF) */
    CIG_CFG_PR_KEEP_IDS = 1u << 9     /* PR: keep the caller's vertex ids inside the engine. By
                                      default a PR engine numbers vertices internally by
                                      decreasing out-degree, so the contributions most gathers
                                      read share cache lines; every id crossing the ABI (records,
                                      gathers, topology) stays the caller's. A renumbered engine
                                      runs no cig_query (CIG_EINVAL). */
};

/* Engine configuration. Flags of the reference samples map as:
 * --cache -> cache_bytes, --n_stream -> n_stream, --SEGMENT -> segments,
 * --block_size -> block_size, --error -> pr_epsilon,
 * --pr_max_rounds -> pr_max_rounds (app_skeleton.h:42-90, hybrid_pr.cu:21-24). */
typedef struct cig_config {
    uint32_t abi_version;     /* must be CIG_ABI_VERSION */
    uint32_t algo;            /* cig_algo */
    uint32_t weight_mode;     /* cig_weight_mode (SSSP only) */
    uint32_t flags;           /* CIG_CFG_* */
    uint64_t weight_seed;     /* seed of CIG_WEIGHT_HASH64 */
    uint64_t cache_bytes;     /* HBM edge-cache budget; 0 = whole graph resident */
    int32_t device;           /* HIP device ordinal */
    uint32_t n_stream;        /* compute streams: 1 = one stream; 2 = a traversal or multi-source
                                 batch groups its records on a second stream beside the
                                 invalidation waves; 3 or
                                 more (default 3) = the deletes also claim their slots there (the
                                 reference spreads its segments' passes over n_stream streams
                                 instead) */
    uint32_t segments;        /* accepted for flag compatibility (1..512) */
    uint32_t block_size;      /* threads per workgroup; 0 = tuned default */
    uint32_t edges_per_lane;  /* edges each lane keeps in flight: 4, 8 or 16 (whole 16-B quads); 0 = tuned default */
    uint32_t split_degree;    /* rows at least this long are split into tiles; 0 = default */
    uint32_t max_rounds;      /* traversal round cap per launch; 0 = unlimited (exact) */
    uint32_t pr_max_rounds;   /* rounds per solve; the reference's flag allows
                                 1..100 (hybrid_pr.cu:22), the library 1..10^6 */
    float pr_epsilon;         /* absolute residual threshold (hybrid_pr.cu:23) */
    uint32_t tile_len;        /* slots per tile of a split row; 0 = default */
    uint32_t round_blocks_per_cu; /* workgroups per CU of round kernels; 0 = default */
    uint32_t pr_pull_divisor; /* PR rounds whose frontier holds at least n / pr_pull_divisor
                                 vertices gather along the incoming rows instead of scattering
                                 (no atomics); 0 = default (4096); >= n: every round */
    uint32_t pull_blocks_per_cu; /* workgroups per CU of the PR pull kernels (at most 128); 0 = default (64) */
    uint32_t kernel_timing_every; /* time one call in this many: its stage markers (the *_ms stage
                                     fields) and, with CIG_CFG_KERNEL_TIMING, its round kernels
                                     (each event record idles the stream a few us); total_ms is
                                     always timed; 0 or 1 = every call */
    uint32_t repair_delta;    /* bucket width of CIG_CFG_ORDERED_REPAIR; 0 = 128 (the reference's) */
    uint32_t pool_slack;      /* edge slots each pool gets beyond its proportional headroom (HBM
                                 without a budget: 1.25x the loaded rows' capacity; host tier:
                                 1.5x), room for row growth before a compaction; 0 = 4 Mi (HBM),
                                 1 Mi (host tier) */
    uint32_t adapt_split;     /* frontier rounds (relax, invalidation) whose waves take fewer than 64
                                 items each raise the split in proportion, keeping rows of up to this
                                 many slots whole in the round instead of cutting them into tiles
                                 expanded a round later; 0 = default (1024); <= split_degree: off */
} cig_config;

/* Per-call statistics. Times are device (HIP event) milliseconds of the
 * stage on the engine's stream; the names follow the reference's
 * [ATTR]/[DELETE-ATTR] timers (hybrid_sssp.cu:511-524, framework.cuh:2874). */
typedef struct cig_batch_stats {
    double total_ms;          /* whole call, submit to results resident */
    double mutate_ms;         /* topology mutation (deletes + adds) */
    double invalidate_ms;     /* deletion seeding + invalidation waves + reset */
    double repair_ms;         /* repair seeding scan */
    double seed_ms;           /* insertion seeding */
    double propagate_ms;      /* fixed-point propagation (relax / PR rounds) */
    double cache_ms;          /* cache admission / eviction / compaction */
    uint64_t edges_scanned;   /* E_scan: edges read by relax/invalidate/repair */
    uint64_t host_edges_scanned; /* of which read zero-copy from the host tier */
    uint64_t vertices_expanded;  /* V_exp: rows expanded (or PR vertices consumed) */
    uint64_t affected;        /* vertices invalidated by deletions */
    uint64_t rounds;          /* fixed-point rounds executed */
    uint64_t deletes_applied; /* delete records that removed an edge */
    uint64_t missing_deletes; /* delete records with no matching edge */
    uint64_t invalid_updates; /* records with an endpoint >= n */
    uint64_t adds_applied;
    uint64_t relocated_rows;  /* rows moved to a larger block (cig_cache_refresh: rows that changed tier) */
    uint64_t pool_compactions;
    uint64_t touched_sources;
    uint32_t converged;       /* PR: 1 if no active residual remains */
    uint32_t launches;        /* kernel launches issued by the call */
    double pr_residual_l1;    /* PR: sum |r| computed in FP64 after the call */
    /* per expansion pass: [0] invalidation waves, [1] repair scan,
     * [2] propagation (relax or PR rounds), [3] PR amendment */
    uint64_t stage_edges[4];
    uint64_t stage_vertices[4];
    uint64_t stage_host_edges[4];
    double scan_kernel_ms;    /* the main repair kernel alone (k_repair_pull, or k_repair_scan
                                 with CIG_CFG_SCAN_REPAIR), 0 if not run */
    double round_kernel_ms;   /* CIG_CFG_KERNEL_TIMING: summed durations of the propagation
                                 round kernels (k_relax_round / k_pr_pull or k_pr_scatter) */
    uint64_t round_launches;  /* round kernels whose durations round_kernel_ms sums (incl. no-op
                                 lookahead); 0 in a call that kernel_timing_every skips */
    uint64_t pull_rounds;     /* PR rounds run in pull form */
    uint64_t inval_rounds;    /* invalidation waves that found work (single-source engines) */
} cig_batch_stats;

typedef struct cig_graph_info {
    uint32_t n;
    uint32_t algo;
    uint64_t edges;             /* live directed edges */
    uint64_t hbm_pool_slots;    /* capacity of the HBM edge cache, in 4-B slots */
    uint64_t hbm_pool_used;     /* bump tail of the HBM pool */
    uint64_t host_pool_slots;   /* capacity of the pinned host tier */
    uint64_t host_pool_used;
    uint64_t cached_rows;       /* rows resident in HBM */
    uint64_t cached_edges;      /* live edges in HBM rows */
    uint64_t device_bytes;      /* total device memory held by the engine */
    uint64_t host_pinned_bytes; /* pinned host memory held by the engine */
    uint32_t grid_blocks;       /* persistent grid size */
    uint32_t block_size;
    uint32_t edges_per_lane;
    uint32_t split_degree;
    uint32_t cu_count;
    uint32_t compactions;         /* pool compactions so far */
    uint32_t inplace_compactions; /* of which packed in place (device kernels for HBM, host
                                     threads for the host tier) */
    uint32_t host_compactions;    /* of which of the host tier */
    uint32_t compaction_ms_hbm;   /* wall time spent compacting each tier, milliseconds */
    uint32_t compaction_ms_host;
    uint32_t refreshes_deferred;  /* serving-loop refreshes that moved nothing because the host tier
                                     would have needed a compaction */
    uint32_t reserved[1];
} cig_graph_info;

/* Result fields for cig_gather. */
typedef enum cig_field {
    CIG_FIELD_VALUE = 0,    /* u32 distance / depth / label; f32 rank for PR */
    CIG_FIELD_PARENT = 1,   /* u32 tight in-neighbour (SSSP/BFS/CC), CIG_INF if none */
    CIG_FIELD_RESIDUAL = 2, /* f32 PR signed residual */
    CIG_FIELD_DEGREE = 3,   /* u32 out-degree */
    CIG_FIELD_TIER = 4,     /* u8: 1 if the row is resident in HBM, 0 if host tier */
    CIG_FIELD_HOTNESS = 5   /* u32 activity score (sum of the 4-window) */
} cig_field;

typedef struct cig_engine cig_engine;

/* ---- lifecycle ---------------------------------------------------------- */
/* Fills *cfg with defaults (SSSP, REF128, whole graph resident, eps 1e-6,
 * 100 PR rounds). */
cig_status cig_config_default(cig_config *cfg);
/* Engine ctor (framework.cuh:545-563). */
cig_status cig_create(const cig_config *cfg, cig_engine **out);
void cig_destroy(cig_engine *e);
/* Message of the last failure on e (or of the last cig_create if e==NULL). */
const char *cig_last_error(const cig_engine *e);

/* ---- graph -------------------------------------------------------------- */
/* Engine::LoadGraph (framework.cuh:800-986): copies a host CSR (offsets has
 * n+1 entries, dst has offsets[n] entries, row order kept) into the
 * two-tier row store. Rows are placed in HBM while the budget lasts
 * (longest rows first), the rest in the pinned host tier. */
cig_status cig_load_csr(cig_engine *e, uint32_t n, const uint64_t *offsets,
                        const uint32_t *dst);
cig_status cig_graph_info_get(cig_engine *e, cig_graph_info *info);
/* Copies the live topology back in CSR form (offsets: n+1 entries; dst
 * capacity dst_capacity). Row order within a row is unspecified when a row
 * received deletes; the multiset equals the reference's. Used by --check. */
cig_status cig_gather_topology(cig_engine *e, uint64_t *offsets, uint32_t *dst,
                               uint64_t dst_capacity);
/* No counterpart in the reference (whose reverse index is a CPU structure,
 * dynamic_reverse_index.h:101-274): copies the device incoming-row index
 * back in CSR form, row v = the sources u of the live edges (u,v), one entry
 * per parallel edge. CIG_ESTATE when the engine keeps no incoming index (PR,
 * or CIG_CFG_SCAN_REPAIR). For verification. */
cig_status cig_gather_incoming(cig_engine *e, uint64_t *offsets, uint32_t *src,
                               uint64_t src_capacity);

/* ---- computation -------------------------------------------------------- */
/* Engine::InitGraph (framework.cuh:991-1041). SSSP/BFS: 1..64 sources, each
 * in [0, n) (the samples clamp theirs, hybrid_sssp.cu:352). With one source
 * the engine keeps one state word per vertex, as the reference. With 2..64
 * sources every later call (cig_converge, the batch calls, the batched form,
 * serving) keeps all of them current together: one mutation per batch, one
 * invalidation wave, repair and relaxation pass for all sources (the reference
 * runs one engine per source); cig_gather then returns nsrc*n words. Several
 * sources need the incoming-row index (not CIG_CFG_SCAN_REPAIR) and chained
 * rounds (not CIG_CFG_COOPERATIVE). CC and PR ignore sources. */
cig_status cig_init(cig_engine *e, const uint32_t *sources, uint32_t nsrc);
/* Engine::Start / StartPageRank (framework.cuh:1151-1182, 1649-1658): runs
 * the fixed point from the initial state. */
cig_status cig_converge(cig_engine *e, cig_batch_stats *stats);
/* Engine::del_edge (framework.cuh:1564-1586 + update_tree_del :2752-2894):
 * applies deletes and repairs the result to the post-delete fixed point. */
cig_status cig_delete_edges(cig_engine *e, const cig_edge *dels, uint64_t n_del,
                            cig_batch_stats *stats);
/* Engine::add_edge (framework.cuh:1397-1401 + update_tree_add :2139-2348). */
cig_status cig_insert_edges(cig_engine *e, const cig_edge *adds, uint64_t n_add,
                            cig_batch_stats *stats);
enum {
    CIG_BATCH_STAGED = 1u << 0,       /* reach the post-delete fixed point before
                                         applying adds (the reference's two
                                         stages); default fuses them into one repair */
    CIG_BATCH_DELETE_PHASE = 1u << 1, /* cig_apply_next_batch: apply only the
                                         current batch's deletes (del_edge); the
                                         batch stays current */
    CIG_BATCH_ADD_PHASE = 1u << 2,    /* cig_apply_next_batch: apply only the
                                         current batch's adds (add_edge) and advance */
    CIG_BATCHES_EACH = 1u << 3        /* cig_apply_next_batches: apply the batches one
                                         at a time, each with its own mutation and fixed
                                         point (every intermediate state is produced);
                                         stats, if not null, points to `count` entries,
                                         one per batch */
};
/* One update batch: all deletes, then all adds (README.md:17-37). For PR this
 * is Engine::UpdatePageRank (framework.cuh:1660-1704). */
cig_status cig_apply_batch(cig_engine *e, const cig_edge *dels, uint64_t n_del,
                           const cig_edge *adds, uint64_t n_add, uint32_t flags,
                           cig_batch_stats *stats);

/* Resident update stream: Engine::get_update_file (framework.cuh:1365-1391)
 * loads every batch up front; batch_sizes holds nbatches (adds, dels) pairs
 * as in the batch-size file (Loader.h:28-61). The records are copied to HBM
 * so that cig_apply_next_batch starts with its inputs resident. */
cig_status cig_load_updates(cig_engine *e, const cig_edge *adds, uint64_t n_add,
                            const cig_edge *dels, uint64_t n_del,
                            const uint32_t *batch_sizes, uint32_t nbatches);
/* Applies the next resident batch (advances the reference's local_begin,
 * framework.cuh:2161-2162). Returns CIG_ELIMIT when no batch is left. */
cig_status cig_apply_next_batch(cig_engine *e, uint32_t flags, cig_batch_stats *stats);
/* ---- serving loop (no counterpart in the reference, whose samples run one
 * batch at a time on the caller's thread) ----------------------------------
 * A server thread owns the engine. Update requests (one batch each) are
 * applied in submission order, queued ones together as one group
 * (cig_apply_next_batches' method). Queries do not hold updates back: they are
 * served when no update waits (or once one has waited max_query_wait_ms),
 * queued queries of one algorithm merged into one multi-source run of up to
 * 64 sources. A query's rounds run on the engine stream against the graph of
 * the version it reports (cig_request_info.version: the update requests
 * applied before it); its results are then copied out on a second stream
 * while the next update group runs. Every result equals the result of the
 * same request served alone at the version it reports. Between groups, when
 * nothing waits, the cache is refreshed every refresh_every groups. A group
 * is only admitted when its staging memory fits in free HBM (hipMemGetInfo)
 * minus hbm_reserve; a request that cannot fit even alone fails with
 * CIG_ENOMEM. While a server runs, the engine may only be used through it. */
typedef struct cig_server cig_server;

typedef struct cig_serve_config {
    uint32_t max_group;          /* update requests per group (default 32) */
    uint32_t max_query_sources;  /* sources merged into one query run (1..64, default 64) */
    uint64_t hbm_reserve;        /* bytes of free HBM admission leaves untouched */
    uint32_t refresh_every;      /* cache refresh (cig_cache_refresh) after this many update groups,
                                    when no request waits; 0 = never; 1 = after every group, as the
                                    reference after every batch (default 8; engines with a budget).
                                    A refresh whose evictions would need a host-tier compaction
                                    moves nothing (cig_graph_info.refreshes_deferred) */
    uint32_t max_query_wait_ms;  /* a query waiting this long is served before further update
                                    groups (default 250) */
    uint64_t reserved[3];
} cig_serve_config;

typedef struct cig_request_info {
    uint32_t status;     /* cig_status of the request (CIG_EPENDING while queued or running) */
    uint32_t kind;       /* 0 update, 1 query */
    uint64_t version;    /* update requests applied before (query) / including (update) it */
    double submit_ms, start_ms, done_ms;  /* server clock: ms since cig_serve_start */
    uint32_t group_size; /* requests served by the same launch chain */
    uint32_t reserved0;
    cig_batch_stats stats; /* of the group or merged query it ran in */
} cig_request_info;

cig_status cig_serve_config_default(cig_serve_config *cfg);
cig_status cig_serve_start(cig_engine *e, const cig_serve_config *cfg, cig_server **out);
/* Copies the records; dels/adds follow cig_apply_batch's semantics. */
cig_status cig_submit_update(cig_server *s, const cig_edge *dels, uint64_t nd, const cig_edge *adds,
                             uint64_t na, uint64_t *ticket);
/* dist (nsrc*n words, source-major as in cig_query) must stay valid until
 * the request is reported done. */
cig_status cig_submit_query(cig_server *s, uint32_t algo, const uint32_t *sources, uint32_t nsrc,
                            uint32_t *dist, uint64_t dist_bytes, uint64_t *ticket);
/* CIG_EPENDING until the request is done; then fills info and forgets the ticket. */
cig_status cig_poll(cig_server *s, uint64_t ticket, cig_request_info *info);
cig_status cig_wait(cig_server *s, uint64_t ticket, cig_request_info *info);
/* Serves what is queued, then stops the thread and frees the server. */
cig_status cig_serve_stop(cig_server *s);

/* Autotuner (no counterpart in the reference, whose launch shape is fixed
 * by --block_size and the CTA scheduler): times the fixed point from
 * `source` on the loaded graph over edges-per-lane x split degree x tile
 * length x relax workgroup size x relax workgroups per CU (coordinate
 * descent, best of two runs each) and keeps the fastest. Results do not
 * depend on the shape. Leaves the engine initialised and converged from
 * `source`. BFS/SSSP/CC engines with chained rounds. */
typedef struct cig_tune_result {
    uint32_t edges_per_lane, split_degree, tile_len, block_size, round_blocks_per_cu, trials;
    double first_ms, best_ms; /* fixed point with the starting and the chosen shape */
    uint64_t reserved[2];
} cig_tune_result;
cig_status cig_autotune(cig_engine *e, uint32_t source, cig_tune_result *result);
/* Multi-source query (no counterpart in the reference's samples): BFS depths
 * or SSSP distances (the engine's weight mode) from up to 64 sources at once
 * on the live graph, both tiers. dist receives nsrc*n words, source-major:
 * dist[s*n + v], UINT32_MAX for unreachable. The engine's own incremental
 * state is not touched. */
cig_status cig_query(cig_engine *e, uint32_t algo, const uint32_t *sources, uint32_t nsrc, uint32_t *dist,
                     uint64_t dist_bytes, cig_batch_stats *stats);
/* No counterpart in the reference. For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): cig_query without any allocation. The run's working memory
 * (vertex-major distances, source masks, stamps: 4*round_up(nsrc,4) + 28
 * bytes per vertex, 36 for BFS) lives in a fixed device array of
 * CIG_NOALLOC_SCRATCH_BYTES that each GPU holds from library load to process
 * exit; sources are passed by value; the distances are written to the
 * caller's device memory dist_device (nsrc*n words, source-major, as
 * cig_query). CIG_ELIMIT when the run needs more than the array. All calls on
 * one GPU share its array: calls that overlap on one GPU (from any engine or
 * thread) race and must not be issued. */
#define CIG_NOALLOC_SCRATCH_BYTES (256ull << 20)
cig_status cig_query_noalloc(cig_engine *e, uint32_t algo, const uint32_t *sources, uint32_t nsrc,
                             uint32_t *dist_device, uint64_t dist_bytes, cig_batch_stats *stats);
/* Batched form (no counterpart in the reference, whose driver applies one
 * batch per call): the next `count` resident batches as one group. Every
 * delete is matched to the latest earlier add of the same (u,v) in the
 * group and the pair cancels; the rest is applied deletes-first with one
 * mutation and one fixed point. Values, parents' validity and the topology
 * multiset after the call equal applying the batches one by one; statistics
 * cover the whole group. Not allowed between the two phases of a batch. */
cig_status cig_apply_next_batches(cig_engine *e, uint32_t count, uint32_t flags, cig_batch_stats *stats);
/* (With CIG_BATCHES_EACH the call is the host loop over the next `count`
 * batches, as the reference's samples run it: no grouping, no state skipped.) */
uint32_t cig_update_batches_left(const cig_engine *e);

/* ---- edge cache (hot rows in HBM, cold rows zero-copy from pinned host) --
 * compute_hot_vertices_* + confirm_candidate_batch + evication_cache +
 * compact_cache + LoadCache (framework.cuh:605-769, 1873-1956): ages the
 * activity window, admits the hottest rows that fit the budget, writes
 * evicted rows back to the host tier and compacts the HBM pool. */
cig_status cig_cache_refresh(cig_engine *e, cig_batch_stats *stats);

/* Packs one tier's edge pool now (tier 0: HBM, 1: pinned host), dropping the
 * garbage that relocated and moved rows left, in place when the live rows fit
 * (else into a larger pool), and empties the tier's free-block lists. No
 * counterpart in the reference, which copies the whole budget at every
 * refresh (driver.cuh:236-247); here packing otherwise happens only when a
 * tier's tail overflows inside a batch, so a caller with idle time (between
 * serving bursts) can take the cost there. Results do not change. */
cig_status cig_compact(cig_engine *e, uint32_t tier);

/* ---- results ------------------------------------------------------------ */
/* GatherValue / GatherParent / GatherBuffer (framework.cuh GraphDatum
 * gathers): copies n elements of the field into out (bytes >= n*elem). An
 * engine initialised with 2..64 sources returns nsrc*n words for VALUE and
 * PARENT, source-major: out[s*n + v] for sources[s]. */
cig_status cig_gather(cig_engine *e, uint32_t field, void *out, size_t bytes);

/* ---- host-side utilities (no GPU needed) -------------------------------- */
/* Deterministic R-MAT generator: 2^scale vertices, edge_factor*2^scale
 * samples, quadrant probabilities a,b,c (d = 1-a-b-c) as in the GAP
 * generator the reference vendors (include/framework/generator.h:80-113),
 * ids not permuted, self-loops and duplicates removed; symmetric != 0 stores
 * each undirected edge both ways. Output arrays are malloc'd by the library
 * (free with cig_host_free). */
cig_status cig_rmat_csr(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                        uint64_t seed, uint32_t symmetric, uint32_t threads,
                        uint32_t *n_out, uint64_t **offsets_out, uint32_t **dst_out);
/* Synthetic update stream over a CSR graph: per batch, n_del deletes sampled
 * uniformly without replacement from the current edges, then n_add R-MAT
 * inserts not present at that point (symmetric: both directions, counts are
 * pairs). Output: adds, dels (malloc'd) and 2*nbatches batch sizes
 * (adds, dels per batch). */
cig_status cig_rmat_stream(uint32_t n, const uint64_t *offsets, const uint32_t *dst,
                           uint32_t scale, double a, double b, double c, uint64_t seed,
                           uint32_t nbatches, uint64_t n_del, uint64_t n_add,
                           uint32_t symmetric,
                           cig_edge **adds_out, uint64_t *n_add_out,
                           cig_edge **dels_out, uint64_t *n_del_out,
                           uint32_t **batch_sizes_out);
/* Text readers: market_big graph (parser.cpp:292-402: "src dst [w]" per line,
 * n = max id + 1) and the update + batch-size files (Loader.h:28-99). */
cig_status cig_read_market(const char *path, uint32_t *n_out, uint64_t **offsets_out,
                           uint32_t **dst_out);
cig_status cig_read_updates(const char *update_path, const char *batch_path,
                            cig_edge **adds_out, uint64_t *n_add_out,
                            cig_edge **dels_out, uint64_t *n_del_out,
                            uint32_t **batch_sizes_out, uint32_t *nbatches_out);
/* Binary forms of the same data (no reference counterpart; the reference
 * reads text only): a 32-byte little-endian header, then the arrays exactly
 * as the C-ABI holds them. Graph "CIGG": n, m, offsets[n+1] (u64), dst[m]
 * (u32). Stream "CIGU": nbatches, n_add, n_del, sizes[2*nbatches] (adds,
 * dels per batch), adds[n_add], dels[n_del] (cig_edge). Readers check the
 * header, monotone offsets, the batch sizes against the record counts and
 * the file length (CIG_EIO otherwise); their arrays are malloc'd (free with
 * cig_host_free). */
cig_status cig_write_csr_bin(const char *path, uint32_t n, const uint64_t *offsets, const uint32_t *dst);
/* Out-of-core R-MAT generator (no counterpart in the reference, whose
 * generator (generator.h:80-113) and loader (csr_graph.cuh:403-479) hold the
 * whole edge list): writes the graph cig_rmat_csr builds from the same
 * arguments to a CIGG file, one range of source segments at a time.
 * `segments` (a power of two <= 2^scale; 0 = 512) are equal vertex ranges;
 * a pass covers consecutive segments whose expected edges fit mem_bytes, so
 * peak memory is about mem_bytes plus one pass's rows, never O(m). *m_out
 * (may be null) receives the edge count. */
cig_status cig_rmat_write(const char *path, uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                          uint64_t seed, uint32_t symmetric, uint32_t segments, uint64_t mem_bytes,
                          uint32_t threads, uint64_t *m_out);
cig_status cig_read_csr_bin(const char *path, uint32_t *n_out, uint64_t **offsets_out, uint32_t **dst_out);
cig_status cig_write_updates_bin(const char *path, const cig_edge *adds, uint64_t n_add,
                                 const cig_edge *dels, uint64_t n_del, const uint32_t *batch_sizes,
                                 uint32_t nbatches);
cig_status cig_read_updates_bin(const char *path, cig_edge **adds_out, uint64_t *n_add_out,
                                cig_edge **dels_out, uint64_t *n_del_out, uint32_t **batch_sizes_out,
                                uint32_t *nbatches_out);
void cig_host_free(void *p);
/* Edge weight of (u,v) under a mode (the function the kernels use). */
uint32_t cig_edge_weight(uint32_t weight_mode, uint64_t seed, uint32_t u, uint32_t v);
/* Number of visible HIP devices (0 when none); never aborts. */
int cig_device_count(void);
/* Diagnostics: times a plain device-to-device copy kernel of `bytes` bytes
 * (16 B per lane) over `iters` launches; *ms_per_launch receives the mean
 * HIP-event time. The HBM reference point for achievable bandwidth and for
 * calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE. No reference counterpart. */
cig_status cig_copy_probe(int32_t device, uint64_t bytes, uint32_t iters, double *ms_per_launch);
/* Diagnostics: times a kernel reading `bytes` bytes of pinned, mapped host
 * memory zero-copy (the access path of rows in the host tier), over `iters`
 * launches. The host-link reference point. No reference counterpart. */
cig_status cig_host_read_probe(int32_t device, uint64_t bytes, uint32_t iters, double *ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif /* CIG_H */
