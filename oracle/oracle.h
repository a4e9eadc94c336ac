/*
 * CPU oracle for the incremental graph engine.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (the coincgraph_amd
 * package, its C-ABI library or the samples) may link, import or call this
 * code. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the checker or as the timed CPU baseline.
 *
 * It is a plain-C restatement of the reference's CPU-side semantics:
 *   - multigraph topology mutation (delete one occurrence per record, ignore
 *     missing deletes, append additions) from
 *     include/groute/graphs/source_local_chunk_store.h:670-848 and the
 *     grouping order of include/framework/effective_update_batch.h:35-87;
 *   - the --check oracles of samples/hybrid_sssp/hybrid_sssp.cu:211-330,
 *     samples/hybrid_bfs/hybrid_bfs.cu (same with unit weights),
 *     samples/hybrid_cc/hybrid_cc_host.cu:8-32 + hybrid_cc.cu:189-201 and
 *     samples/hybrid_pr/hybrid_pr_host.cu:9-56;
 *   - full recomputes: Dijkstra, queue BFS, the CCCheck flood fill, FP64
 *     Jacobi PageRank, and the label-correcting solver of
 *     samples/hybrid_sssp/hybrid_sssp_host.cu:34-63 (with the engine's +1).
 *
 * Every function is pinned by tests/test_oracle.py against an independent
 * route (scipy.sparse.csgraph, numpy.linalg.solve, brute force, hand-derived
 * known answers); see the test docstrings for which check pins what.
 */
#ifndef CIG_ORACLE_H
#define CIG_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OG_INF 0xFFFFFFFFu

/* weight modes (match include/cig.h cig_weight_mode, plus unit/zero) */
enum { OG_W_REF128 = 0, OG_W_HASH64 = 1, OG_W_UNIT = 2, OG_W_ZERO = 3 };

typedef struct og_graph {
    uint32_t n;
    uint32_t *deg;
    uint32_t *cap;
    uint32_t **rows;
} og_graph;

/* Edge weight of (u,v). REF128 is api.cuh:26-28 / hybrid_sssp.cu:236
 * ((u+v)%128+1 in 64-bit); HASH64 is 1 + splitmix64(seed ^ (u<<32|v)) % 64. */
uint32_t og_weight(int mode, uint64_t seed, uint32_t u, uint32_t v);
uint64_t og_splitmix64(uint64_t x);

og_graph *og_create(uint32_t n, const uint64_t *offsets, const uint32_t *dst);
void og_free(og_graph *g);
uint64_t og_edge_count(const og_graph *g);
/* CSR export in stored row order; offsets has n+1 entries. */
void og_export(const og_graph *g, uint64_t *offsets, uint32_t *dst);

/* Reference mutation semantics. Records with an endpoint >= n are counted in
 * *invalid and ignored (the reference counts source >= n the same way,
 * source_local_chunk_store.h:258-264; it would append an out-of-range
 * destination, which the engine rejects instead). Deletes remove, per source,
 * the first min(requested, present) occurrences of each destination in row
 * order, keeping survivor order (BuildDeletionPlan + CompactSourceInPlace,
 * :670-780); unmatched deletes are counted in *missing. Adds are appended per
 * source in input order (stable grouping, effective_update_batch.h:49-71). */
int og_apply_deletes(og_graph *g, const uint32_t *u, const uint32_t *v, uint64_t k,
                     uint64_t *missing, uint64_t *invalid);
int og_apply_adds(og_graph *g, const uint32_t *u, const uint32_t *v, uint64_t k,
                  uint64_t *invalid);
/* Deletes then adds with `threads` workers over source ranges (the
 * reference's CPU mutation pool, source_local_chunk_store.h:116-121, 218-391);
 * same result as og_apply_deletes followed by og_apply_adds. */
int og_apply_batch_mt(og_graph *g, const uint32_t *du, const uint32_t *dv, uint64_t nd, const uint32_t *au,
                      const uint32_t *av, uint64_t na, uint32_t threads, uint64_t *missing,
                      uint64_t *invalid);

/* Full recomputes. */
void og_sssp_dijkstra(const og_graph *g, uint32_t src, int wmode, uint64_t seed,
                      uint32_t *dist);
/* Returns the number of edges examined (the CPU baseline's E_scan). */
uint64_t og_sssp_label_correcting(const og_graph *g, uint32_t src, int wmode, uint64_t seed,
                                  uint32_t *dist);
void og_bfs(const og_graph *g, uint32_t src, uint32_t *depth);
void og_cc_flood(const og_graph *g, uint32_t *labels);
/* FP64 Jacobi from x=0 until sum|x'-x| <= tol_per_vertex * max(1,n) or
 * max_rounds; returns rounds used (negative if not converged). */
int og_pr_jacobi(const og_graph *g, double *x, double tol_per_vertex, int max_rounds);

/* --check oracles. */
typedef struct {
    uint64_t relaxable_edges;
    uint64_t missing_tight_witnesses;
    uint32_t first_src, first_dst, first_missing_witness, pad;
} og_bellman_result;

typedef struct {
    uint64_t invalid_vertices;
    uint64_t missing_parent_edges;
    uint64_t wrong_parent_distances;
    uint32_t first_vertex, first_parent;
} og_parent_result;

typedef struct {
    double l1_error, bound, max_error, invariant_l1, max_residual, reference_mass;
    int converged, finite, ok, pad;
} og_pr_check_result;

void og_check_bellman(const og_graph *g, const uint32_t *dist, uint32_t src, int wmode,
                      uint64_t seed, og_bellman_result *out);
void og_check_parent(const og_graph *g, const uint32_t *dist, const uint32_t *parent,
                     uint32_t src, int wmode, uint64_t seed, og_parent_result *out);
uint64_t og_distance_checksum(const uint32_t *dist, uint32_t n);
uint64_t og_parent_checksum(const uint32_t *dist, const uint32_t *parent, uint32_t n);
uint64_t og_cc_check(const og_graph *g, const uint32_t *labels);
void og_pr_check(const og_graph *g, const float *ranks, const float *residual,
                 double epsilon, og_pr_check_result *out);
/* The invariant half of PageRankCheck (hybrid_pr_host.cu:15-45) without the
 * reference solve, for graphs too large for a Jacobi run per batch:
 * defect = sum_v |0.15 + 0.85*sum_{u->v} x[u]/outdeg(u) - x[v] - r[v]|,
 * max_res = max|r|, mass = sum|x|; returns 0 if any x or r is not finite. */
int og_pr_invariant(const og_graph *g, const float *ranks, const float *residual,
                    double *defect, double *max_res, double *mass);

#ifdef __cplusplus
}
#endif
#endif
