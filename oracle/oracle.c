/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h). Plain C99.
 * Each function cites the reference file:line whose behaviour it restates.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

uint64_t og_splitmix64(uint64_t x) {
    /* Vigna's splitmix64 finaliser (public-domain published constants). */
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

uint32_t og_weight(int mode, uint64_t seed, uint32_t u, uint32_t v) {
    switch (mode) {
    case OG_W_REF128: /* api.cuh:26-28, hybrid_sssp.cu:236 */
        return (uint32_t)(((uint64_t)u + (uint64_t)v) % 128u + 1u);
    case OG_W_HASH64:
        return (uint32_t)(1u + og_splitmix64(seed ^ (((uint64_t)u << 32) | v)) % 64u);
    case OG_W_UNIT: /* hybrid_bfs.cu: TraversalEdgeWeight = 1 */
        return 1u;
    default: /* hybrid_cc.cu:55-57: zero traversal cost */
        return 0u;
    }
}

static int row_reserve(og_graph *g, uint32_t s, uint64_t need) {
    if (need <= g->cap[s]) return 0;
    uint64_t c = g->cap[s] ? g->cap[s] : 4;
    while (c < need) c *= 2;
    uint32_t *p = (uint32_t *)realloc(g->rows[s], c * sizeof(uint32_t));
    if (!p) return -1;
    g->rows[s] = p;
    g->cap[s] = (uint32_t)c;
    return 0;
}

og_graph *og_create(uint32_t n, const uint64_t *offsets, const uint32_t *dst) {
    og_graph *g = (og_graph *)calloc(1, sizeof(og_graph));
    if (!g) return NULL;
    g->n = n;
    g->deg = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    g->cap = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    g->rows = (uint32_t **)calloc(n ? n : 1, sizeof(uint32_t *));
    if (!g->deg || !g->cap || !g->rows) { og_free(g); return NULL; }
    for (uint32_t s = 0; s < n; ++s) {
        uint64_t d = offsets[s + 1] - offsets[s];
        if (d == 0) continue;
        if (row_reserve(g, s, d)) { og_free(g); return NULL; }
        memcpy(g->rows[s], dst + offsets[s], d * sizeof(uint32_t));
        g->deg[s] = (uint32_t)d;
    }
    return g;
}

void og_free(og_graph *g) {
    if (!g) return;
    if (g->rows)
        for (uint32_t s = 0; s < g->n; ++s) free(g->rows[s]);
    free(g->rows);
    free(g->deg);
    free(g->cap);
    free(g);
}

uint64_t og_edge_count(const og_graph *g) {
    uint64_t m = 0;
    for (uint32_t s = 0; s < g->n; ++s) m += g->deg[s];
    return m;
}

void og_export(const og_graph *g, uint64_t *offsets, uint32_t *dst) {
    uint64_t o = 0;
    for (uint32_t s = 0; s < g->n; ++s) {
        offsets[s] = o;
        if (g->deg[s]) memcpy(dst + o, g->rows[s], (size_t)g->deg[s] * sizeof(uint32_t));
        o += g->deg[s];
    }
    offsets[g->n] = o;
}

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* Removes from row s the first occurrences of the len requested
 * destinations list[] (sorted in place; rem is scratch of len words); returns
 * the number of requests that matched an edge. */
static uint64_t delete_from_row(og_graph *g, uint32_t s, uint32_t *list, uint64_t len, uint32_t *rem) {
    qsort(list, len, sizeof(uint32_t), cmp_u32);
    /* run-length: unique destinations in list[0..nu), multiplicity in rem */
    uint64_t nu = 0;
    for (uint64_t i = 0; i < len; ++i) {
        if (nu && list[nu - 1] == list[i]) { rem[nu - 1]++; continue; }
        list[nu] = list[i];
        rem[nu] = 1;
        ++nu;
    }
    uint64_t matched = 0, w = 0;
    uint32_t *row = g->rows[s];
    for (uint64_t r = 0; r < g->deg[s]; ++r) {
        uint32_t d = row[r];
        uint64_t lo = 0, hi = nu;
        while (lo < hi) {
            uint64_t mid = (lo + hi) / 2;
            if (list[mid] < d) lo = mid + 1; else hi = mid;
        }
        if (lo < nu && list[lo] == d && rem[lo] > 0) {
            rem[lo]--;
            ++matched;
            continue;
        }
        row[w++] = d;
    }
    g->deg[s] = (uint32_t)w;
    return matched;
}

/* Deletion phase: BuildDeletionPlan (source_local_chunk_store.h:670-723)
 * collects per source the requested multiplicity of each destination; the
 * compaction (:733-780) then removes, scanning the row in order, the first
 * min(requested, present) occurrences of each destination and keeps the
 * surviving order. */
int og_apply_deletes(og_graph *g, const uint32_t *u, const uint32_t *v, uint64_t k,
                     uint64_t *missing, uint64_t *invalid) {
    uint64_t miss = 0, inval = 0;
    uint32_t n = g->n;
    uint64_t *count = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    if (!count) return -1;
    for (uint64_t i = 0; i < k; ++i) {
        if (u[i] >= n || v[i] >= n) { ++inval; continue; }
        count[u[i] + 1]++;
    }
    for (uint32_t s = 0; s < n; ++s) count[s + 1] += count[s];
    uint64_t total = count[n];
    uint32_t *req = (uint32_t *)malloc((total ? total : 1) * sizeof(uint32_t));
    uint64_t *fill = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
    uint32_t *remaining = (uint32_t *)malloc((total ? total : 1) * sizeof(uint32_t));
    if (!req || !fill || !remaining) { free(count); free(req); free(fill); free(remaining); return -1; }
    memcpy(fill, count, ((size_t)n + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < k; ++i) {
        if (u[i] >= n || v[i] >= n) continue;
        req[fill[u[i]]++] = v[i];
    }
    for (uint32_t s = 0; s < n; ++s) {
        uint64_t b = count[s], e = count[s + 1];
        if (b == e) continue;
        miss += (e - b) - delete_from_row(g, s, req + b, e - b, remaining + b);
    }
    free(count); free(req); free(fill); free(remaining);
    if (missing) *missing += miss;
    if (invalid) *invalid += inval;
    return 0;
}

/* Addition phase: appended per source in input order
 * (RewriteSourceToNewBlock / CompactSourceInPlace append, :764-770, :827-833). */
int og_apply_adds(og_graph *g, const uint32_t *u, const uint32_t *v, uint64_t k,
                  uint64_t *invalid) {
    uint64_t inval = 0;
    for (uint64_t i = 0; i < k; ++i) {
        if (u[i] >= g->n || v[i] >= g->n) { ++inval; continue; }
        uint32_t s = u[i];
        if (row_reserve(g, s, (uint64_t)g->deg[s] + 1)) return -1;
        g->rows[s][g->deg[s]++] = v[i];
    }
    if (invalid) *invalid += inval;
    return 0;
}

/* The reference's CPU batch maintenance with a worker pool
 * (ApplyMutationPhase, source_local_chunk_store.h:116-121, 218-391): the
 * records are grouped by source (stable, as GroupedUpdateBatch,
 * effective_update_batch.h:35-87), then `threads` workers each take a range
 * of sources with about equal record counts and apply, per source, its
 * deletes (delete_from_row) and then its adds in input order. The result is
 * the same as og_apply_deletes followed by og_apply_adds. */
typedef struct {
    og_graph *g;
    const uint64_t *dcnt, *acnt;
    uint32_t *dreq, *drem;
    const uint32_t *areq;
    uint32_t lo, hi;
    uint64_t matched;
    int err;
} mt_job;

static void *mt_worker(void *arg) {
    mt_job *j = (mt_job *)arg;
    for (uint32_t s = j->lo; s < j->hi; ++s) {
        uint64_t b = j->dcnt[s], e = j->dcnt[s + 1];
        if (b < e) j->matched += delete_from_row(j->g, s, j->dreq + b, e - b, j->drem + b);
        b = j->acnt[s];
        e = j->acnt[s + 1];
        if (b == e) continue;
        if (row_reserve(j->g, s, (uint64_t)j->g->deg[s] + (e - b))) { j->err = 1; return NULL; }
        memcpy(j->g->rows[s] + j->g->deg[s], j->areq + b, (e - b) * sizeof(uint32_t));
        j->g->deg[s] += (uint32_t)(e - b);
    }
    return NULL;
}

/* per-source counts (n+1 prefix) and the destinations grouped by source */
static int group_by_source(uint32_t n, const uint32_t *u, const uint32_t *v, uint64_t k, uint64_t **cnt_out,
                           uint32_t **req_out, uint64_t *inval) {
    uint64_t *cnt = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    if (!cnt) return -1;
    for (uint64_t i = 0; i < k; ++i) {
        if (u[i] >= n || v[i] >= n) { ++*inval; continue; }
        cnt[u[i] + 1]++;
    }
    for (uint32_t s = 0; s < n; ++s) cnt[s + 1] += cnt[s];
    uint64_t *fill = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
    uint32_t *req = (uint32_t *)malloc((cnt[n] ? cnt[n] : 1) * sizeof(uint32_t));
    if (!fill || !req) { free(cnt); free(fill); free(req); return -1; }
    memcpy(fill, cnt, ((size_t)n + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < k; ++i)
        if (u[i] < n && v[i] < n) req[fill[u[i]]++] = v[i];
    free(fill);
    *cnt_out = cnt;
    *req_out = req;
    return 0;
}

int og_apply_batch_mt(og_graph *g, const uint32_t *du, const uint32_t *dv, uint64_t nd, const uint32_t *au,
                      const uint32_t *av, uint64_t na, uint32_t threads, uint64_t *missing,
                      uint64_t *invalid) {
    uint32_t n = g->n;
    uint64_t inval = 0, *dcnt = NULL, *acnt = NULL;
    uint32_t *dreq = NULL, *areq = NULL;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if (group_by_source(n, du, dv, nd, &dcnt, &dreq, &inval) ||
        group_by_source(n, au, av, na, &acnt, &areq, &inval)) {
        free(dcnt); free(dreq); free(acnt); free(areq);
        return -1;
    }
    uint32_t *drem = (uint32_t *)malloc((dcnt[n] ? dcnt[n] : 1) * sizeof(uint32_t));
    mt_job *jobs = (mt_job *)calloc(threads, sizeof(mt_job));
    pthread_t *tid = (pthread_t *)calloc(threads, sizeof(pthread_t));
    int rc = (!drem || !jobs || !tid) ? -1 : 0;
    if (rc == 0) {
        /* source ranges of about equal record counts */
        uint64_t total = dcnt[n] + acnt[n], s = 0;
        for (uint32_t t = 0; t < threads; ++t) {
            uint64_t goal = total * (t + 1) / threads;
            jobs[t].lo = (uint32_t)s;
            while (s < n && (t + 1 == threads || dcnt[s + 1] + acnt[s + 1] <= goal)) ++s;
            jobs[t].hi = (uint32_t)s;
            jobs[t].g = g;
            jobs[t].dcnt = dcnt;
            jobs[t].acnt = acnt;
            jobs[t].dreq = dreq;
            jobs[t].drem = drem;
            jobs[t].areq = areq;
        }
        uint32_t started = 0;
        for (; started < threads; ++started)
            if (pthread_create(&tid[started], NULL, mt_worker, &jobs[started]) != 0) break;
        for (uint32_t t = started; t < threads; ++t) mt_worker(&jobs[t]); /* could not start: run here */
        for (uint32_t t = 0; t < started; ++t) pthread_join(tid[t], NULL);
        uint64_t matched = 0;
        for (uint32_t t = 0; t < threads; ++t) {
            matched += jobs[t].matched;
            if (jobs[t].err) rc = -1;
        }
        if (missing) *missing += dcnt[n] - matched;
        if (invalid) *invalid += inval;
    }
    free(dcnt); free(dreq); free(acnt); free(areq); free(drem); free(jobs); free(tid);
    return rc;
}

/* ---- binary min-heap on (dist, vertex) for Dijkstra ---- */
typedef struct { uint64_t d; uint32_t v; } hent;

static void heap_push(hent *h, uint64_t *sz, hent x) {
    uint64_t i = (*sz)++;
    while (i) {
        uint64_t p = (i - 1) / 2;
        if (h[p].d <= x.d) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = x;
}

static hent heap_pop(hent *h, uint64_t *sz) {
    hent top = h[0], x = h[--(*sz)];
    uint64_t i = 0, n = *sz;
    for (;;) {
        uint64_t l = 2 * i + 1, r = l + 1, m = i;
        hent best = x;
        if (l < n && h[l].d < best.d) { m = l; best = h[l]; }
        if (r < n && h[r].d < best.d) { m = r; best = h[r]; }
        if (m == i) break;
        h[i] = h[m];
        i = m;
    }
    if (n) h[i] = x;
    return top;
}

void og_sssp_dijkstra(const og_graph *g, uint32_t src, int wmode, uint64_t seed,
                      uint32_t *dist) {
    uint32_t n = g->n;
    for (uint32_t i = 0; i < n; ++i) dist[i] = OG_INF;
    if (src >= n) return;
    uint64_t m = og_edge_count(g);
    hent *h = (hent *)malloc((m + 2) * sizeof(hent));
    uint64_t *best = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    for (uint32_t i = 0; i < n; ++i) best[i] = UINT64_MAX;
    uint64_t sz = 0;
    best[src] = 0;
    heap_push(h, &sz, (hent){0, src});
    while (sz) {
        hent t = heap_pop(h, &sz);
        if (t.d != best[t.v]) continue;
        const uint32_t *row = g->rows[t.v];
        for (uint32_t e = 0; e < g->deg[t.v]; ++e) {
            uint32_t d = row[e];
            uint64_t c = t.d + og_weight(wmode, seed, t.v, d);
            if (c < best[d]) {
                best[d] = c;
                heap_push(h, &sz, (hent){c, d});
            }
        }
    }
    for (uint32_t i = 0; i < n; ++i) dist[i] = best[i] >= OG_INF ? OG_INF : (uint32_t)best[i];
    free(h);
    free(best);
}

/* samples/hybrid_sssp/hybrid_sssp_host.cu:34-63 (FIFO label-correcting),
 * with the engine's weight function (that file's %128 without +1 disagrees
 * with the engine and the checker; the engine's is used). */
uint64_t og_sssp_label_correcting(const og_graph *g, uint32_t src, int wmode, uint64_t seed,
                                  uint32_t *dist) {
    uint32_t n = g->n;
    uint64_t scanned = 0;
    for (uint32_t i = 0; i < n; ++i) dist[i] = OG_INF;
    if (src >= n) return 0;
    /* ring-buffer FIFO; a vertex may be queued more than once, so bound the
     * live size by n+1 using an in-queue flag (same fixed point). */
    uint32_t *q = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    uint8_t *inq = (uint8_t *)calloc(n, 1);
    uint64_t head = 0, tail = 0, cap = (uint64_t)n + 1;
    dist[src] = 0;
    q[tail++ % cap] = src;
    inq[src] = 1;
    while (head != tail) {
        uint32_t x = q[head++ % cap];
        inq[x] = 0;
        uint64_t dx = dist[x];
        scanned += g->deg[x];
        for (uint32_t e = 0; e < g->deg[x]; ++e) {
            uint32_t d = g->rows[x][e];
            uint64_t c = dx + og_weight(wmode, seed, x, d);
            if (c < dist[d]) {
                dist[d] = (uint32_t)c;
                if (!inq[d]) { inq[d] = 1; q[tail++ % cap] = d; }
            }
        }
    }
    free(q);
    free(inq);
    return scanned;
}

void og_bfs(const og_graph *g, uint32_t src, uint32_t *depth) {
    uint32_t n = g->n;
    for (uint32_t i = 0; i < n; ++i) depth[i] = OG_INF;
    if (src >= n) return;
    uint32_t *q = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint64_t head = 0, tail = 0;
    depth[src] = 0;
    q[tail++] = src;
    while (head < tail) {
        uint32_t x = q[head++];
        for (uint32_t e = 0; e < g->deg[x]; ++e) {
            uint32_t d = g->rows[x][e];
            if (depth[d] == OG_INF) { depth[d] = depth[x] + 1; q[tail++] = d; }
        }
    }
    free(q);
}

/* samples/hybrid_cc/hybrid_cc_host.cu:8-32: flood from seeds in ascending ID
 * order; an already-labelled vertex (and what it reaches) has a smaller seed. */
void og_cc_flood(const og_graph *g, uint32_t *labels) {
    uint32_t n = g->n;
    for (uint32_t i = 0; i < n; ++i) labels[i] = OG_INF;
    uint32_t *stack = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    for (uint32_t seed = 0; seed < n; ++seed) {
        if (labels[seed] != OG_INF) continue;
        uint64_t sp = 0;
        labels[seed] = seed;
        stack[sp++] = seed;
        while (sp) {
            uint32_t x = stack[--sp];
            for (uint32_t e = 0; e < g->deg[x]; ++e) {
                uint32_t d = g->rows[x][e];
                if (labels[d] != OG_INF) continue;
                labels[d] = seed;
                stack[sp++] = d;
            }
        }
    }
    free(stack);
}

uint64_t og_cc_check(const og_graph *g, const uint32_t *labels) {
    uint32_t *expected = (uint32_t *)malloc(((size_t)g->n + 1) * sizeof(uint32_t));
    og_cc_flood(g, expected);
    uint64_t errors = 0;
    for (uint32_t i = 0; i < g->n; ++i) errors += labels[i] != expected[i];
    free(expected);
    return errors;
}

/* PageRankCheck's multiply (hybrid_pr_host.cu:15-24):
 * next = 0.15 + sum over edges u->v of 0.85*x[u]/outdeg(u); dangling drop. */
static void pr_multiply(const og_graph *g, const double *x, double *next) {
    for (uint32_t v = 0; v < g->n; ++v) next[v] = 0.15;
    for (uint32_t s = 0; s < g->n; ++s) {
        if (!g->deg[s]) continue;
        double c = 0.85 * x[s] / g->deg[s];
        for (uint32_t e = 0; e < g->deg[s]; ++e) next[g->rows[s][e]] += c;
    }
}

int og_pr_jacobi(const og_graph *g, double *x, double tol_per_vertex, int max_rounds) {
    uint32_t n = g->n;
    double *next = (double *)malloc(((size_t)n + 1) * sizeof(double));
    for (uint32_t v = 0; v < n; ++v) x[v] = 0;
    int rounds = -max_rounds;
    for (int r = 0; r < max_rounds; ++r) {
        pr_multiply(g, x, next);
        double delta = 0;
        for (uint32_t v = 0; v < n; ++v) delta += fabs(next[v] - x[v]);
        memcpy(x, next, (size_t)n * sizeof(double));
        if (delta <= tol_per_vertex * (double)(n ? n : 1)) { rounds = r + 1; break; }
    }
    free(next);
    return rounds;
}

/* hybrid_pr_host.cu:9-56, restated. */
void og_pr_check(const og_graph *g, const float *ranks, const float *residual,
                 double epsilon, og_pr_check_result *out) {
    uint32_t n = g->n;
    double *ref = (double *)malloc(((size_t)n + 1) * sizeof(double));
    double *next = (double *)malloc(((size_t)n + 1) * sizeof(double));
    double *actual = (double *)calloc((size_t)n + 1, sizeof(double));
    int r = og_pr_jacobi(g, ref, 1e-12, 10000);
    for (uint32_t v = 0; v < n; ++v) actual[v] = ranks[v];
    pr_multiply(g, actual, next);
    double l1 = 0, defect = 0, max_res = 0, max_err = 0, mass = 0;
    int finite = 1;
    for (uint32_t v = 0; v < n; ++v) {
        finite &= isfinite(ranks[v]) && isfinite(residual[v]) && ranks[v] >= -epsilon;
        double err = fabs((double)ranks[v] - ref[v]);
        l1 += err;
        if (err > max_err) max_err = err;
        defect += fabs(next[v] - (double)ranks[v] - (double)residual[v]);
        if (fabs((double)residual[v]) > max_res) max_res = fabs((double)residual[v]);
        mass += fabs(ref[v]);
    }
    double roundoff = 64.0 * FLT_EPSILON * (mass > 1.0 ? mass : 1.0);
    double bound = ((double)n * epsilon + roundoff) / 0.15;
    out->l1_error = l1;
    out->bound = bound;
    out->max_error = max_err;
    out->invariant_l1 = defect;
    out->max_residual = max_res;
    out->reference_mass = mass;
    out->converged = r > 0;
    out->finite = finite;
    out->ok = finite && r > 0 && l1 <= bound && defect <= roundoff &&
              max_res <= (double)(float)epsilon;
    free(ref);
    free(next);
    free(actual);
}

/* hybrid_pr_host.cu:15-45: the residual invariant x + r = 0.15 + 0.85*P^T x
 * and the largest residual, with no reference solve. */
int og_pr_invariant(const og_graph *g, const float *ranks, const float *residual,
                    double *defect, double *max_res, double *mass) {
    uint32_t n = g->n;
    double *x = (double *)calloc((size_t)n + 1, sizeof(double));
    double *next = (double *)calloc((size_t)n + 1, sizeof(double));
    int finite = 1;
    double d = 0, mr = 0, m = 0;
    for (uint32_t v = 0; v < n; ++v) x[v] = ranks[v];
    pr_multiply(g, x, next);
    for (uint32_t v = 0; v < n; ++v) {
        finite &= isfinite(ranks[v]) && isfinite(residual[v]);
        d += fabs(next[v] - (double)ranks[v] - (double)residual[v]);
        if (fabs((double)residual[v]) > mr) mr = fabs((double)residual[v]);
        m += fabs((double)ranks[v]);
    }
    *defect = d;
    *max_res = mr;
    *mass = m;
    free(x);
    free(next);
    return finite;
}

/* hybrid_sssp.cu:211-264 (CheckBellmanRelaxed). */
void og_check_bellman(const og_graph *g, const uint32_t *dist, uint32_t src, int wmode,
                      uint64_t seed, og_bellman_result *out) {
    uint32_t n = g->n;
    memset(out, 0, sizeof(*out));
    out->first_src = out->first_dst = out->first_missing_witness = OG_INF;
    uint8_t *tight = (uint8_t *)calloc(n ? n : 1, 1);
    if (src < n) tight[src] = 1;
    for (uint32_t s = 0; s < n; ++s) {
        if (dist[s] == OG_INF) continue;
        for (uint32_t e = 0; e < g->deg[s]; ++e) {
            uint32_t d = g->rows[s][e];
            if (d >= n) continue;
            uint64_t c = (uint64_t)dist[s] + og_weight(wmode, seed, s, d);
            if (dist[d] != OG_INF && c == dist[d]) tight[d] = 1;
            if (c < dist[d]) {
                if (out->relaxable_edges == 0) { out->first_src = s; out->first_dst = d; }
                out->relaxable_edges++;
            }
        }
    }
    for (uint32_t v = 0; v < n; ++v) {
        if (dist[v] == OG_INF || tight[v]) continue;
        if (out->missing_tight_witnesses == 0) out->first_missing_witness = v;
        out->missing_tight_witnesses++;
    }
    free(tight);
}

/* Open-addressing set of (parent, vertex) pairs: CheckParentWitness asks,
 * for every reached vertex, whether its parent's row holds it. Looking each
 * pair up while scanning every row once costs O(E + n) instead of
 * O(sum of the parents' degrees), which hub parents make quadratic. */
typedef struct {
    uint64_t *key; /* p << 32 | v, or ~0 for an empty slot */
    uint8_t *hit;
    uint64_t mask;
} pair_set;

static uint64_t pair_hash(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return k;
}

static int pair_set_init(pair_set *t, uint64_t count) {
    uint64_t cap = 16;
    while (cap < 2 * count) cap <<= 1;
    t->mask = cap - 1;
    t->key = (uint64_t *)malloc(cap * sizeof(uint64_t));
    t->hit = (uint8_t *)calloc(cap, 1);
    if (!t->key || !t->hit) return -1;
    memset(t->key, 0xFF, cap * sizeof(uint64_t));
    return 0;
}

/* slot of k; inserted when absent and insert != 0, else ~0 when absent */
static uint64_t pair_set_slot(pair_set *t, uint64_t k, int insert) {
    for (uint64_t i = pair_hash(k) & t->mask;; i = (i + 1) & t->mask) {
        if (t->key[i] == k) return i;
        if (t->key[i] == ~0ull) {
            if (!insert) return ~0ull;
            t->key[i] = k;
            return i;
        }
    }
}

/* hybrid_sssp.cu:266-316 (CheckParentWitness). */
void og_check_parent(const og_graph *g, const uint32_t *dist, const uint32_t *parent,
                     uint32_t src, int wmode, uint64_t seed, og_parent_result *out) {
    uint32_t n = g->n;
    memset(out, 0, sizeof(*out));
    out->first_vertex = out->first_parent = OG_INF;
    pair_set set;
    if (pair_set_init(&set, n) != 0) {
        free(set.key);
        free(set.hit);
        out->invalid_vertices = n;
        return;
    }
    for (uint32_t v = 0; v < n; ++v) {
        if (v == src || dist[v] == OG_INF) continue;
        uint32_t p = parent[v];
        if (p < n && dist[p] != OG_INF) pair_set_slot(&set, (uint64_t)p << 32 | v, 1);
    }
    for (uint32_t s = 0; s < n; ++s)
        for (uint32_t e = 0; e < g->deg[s]; ++e) {
            uint64_t i = pair_set_slot(&set, (uint64_t)s << 32 | g->rows[s][e], 0);
            if (i != ~0ull) set.hit[i] = 1;
        }
    for (uint32_t v = 0; v < n; ++v) {
        if (v == src || dist[v] == OG_INF) continue;
        uint32_t p = parent[v];
        int valid = p < n && dist[p] != OG_INF;
        if (valid) {
            int exists = set.hit[pair_set_slot(&set, (uint64_t)p << 32 | v, 0)];
            int matches = (uint64_t)dist[p] + og_weight(wmode, seed, p, v) == dist[v];
            if (!exists) out->missing_parent_edges++;
            else if (!matches) out->wrong_parent_distances++;
            valid = exists && matches;
        }
        if (!valid) {
            if (out->invalid_vertices == 0) { out->first_vertex = v; out->first_parent = p; }
            out->invalid_vertices++;
        }
    }
    free(set.key);
    free(set.hit);
}

/* hybrid_sssp.cu:318-330 (DistanceChecksum); hybrid_cc.cu:189-201 is the
 * same formula over labels. */
uint64_t og_distance_checksum(const uint32_t *dist, uint32_t n) {
    uint64_t c = 1469598103934665603ull;
    for (uint32_t v = 0; v < n; ++v) {
        if (dist[v] == OG_INF) continue;
        c ^= (uint64_t)v + 0x9e3779b97f4a7c15ull + ((uint64_t)dist[v] << 6) +
             ((uint64_t)dist[v] >> 2);
        c *= 1099511628211ull;
    }
    return c;
}

/* hybrid_sssp.cu:604-619 (parent checksum printed with --sssp_print_checksum). */
uint64_t og_parent_checksum(const uint32_t *dist, const uint32_t *parent, uint32_t n) {
    uint64_t c = 1099511628211ull;
    for (uint32_t v = 0; v < n; ++v) {
        if (dist[v] == OG_INF) continue;
        c ^= (uint64_t)v + 0x517cc1b727220a95ull + ((uint64_t)parent[v] << 7) +
             ((uint64_t)parent[v] >> 3);
        c *= 1469598103934665603ull;
    }
    return c;
}
