// Kernels of the incremental graph engine. Included once, by engine.hip.
#pragma once
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "../../include/cig.h"
#include "device.hpp"

namespace cig {
namespace dev {

namespace cg = cooperative_groups;

__device__ __forceinline__ uint64_t next_pow2_dev(uint64_t d) {
    uint64_t c = d < kMinRowCap ? kMinRowCap : d;
    return c <= 1 ? 1 : (1ull << (64 - __builtin_clzll(c - 1)));
}

__device__ __forceinline__ uint32_t gthread() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint32_t gthreads() { return gridDim.x * blockDim.x; }
__device__ __forceinline__ uint32_t gwave() { return blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; }
__device__ __forceinline__ uint32_t gwaves() { return gridDim.x * (blockDim.x / kWave); }

// ------------------------------------------------------------- small kernels
// Zero-fills a 32-bit range (used instead of hipMemsetAsync so the batch chain
// stays capturable in a HIP graph).
__global__ void k_fill32(uint32_t *p, uint64_t n, uint32_t value) {
    for (uint64_t i = gthread(); i < n; i += gthreads()) p[i] = value;
}

// Per-call reset of the control block counters (not the round state).
__global__ void k_reset_ctl(Ctl *c, uint32_t keep_rounds) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c->acount = 0;
    c->icount[0] = c->icount[1] = c->icount[2] = 0;
    c->scount = 0;
    c->touched = 0;
    c->small_cnt = c->big_cnt = c->add_cnt = c->copy_cnt = c->overflow_cnt = 0;
    c->dtile_cnt = c->ht_cursor = 0;
    c->pr_invalid = 0;
    c->group_cursor = 0;
    c->edges_scanned = c->host_edges = c->vexp = 0;
    c->deletes_applied = c->missing = c->invalid = c->adds_applied = c->relocated = 0;
    c->rounds = 0;
    c->residual_l1 = 0;
    c->converged = 0;
    for (int i = 0; i < 4; ++i) c->st_edges[i] = c->st_vexp[i] = c->st_host[i] = 0;
    if (!keep_rounds) {
        c->round = 1;
        for (int i = 0; i < 3; ++i) c->qcount[i] = c->tcount[i] = 0;
    }
}

// Clears the per-mutation lists between the two stages of a staged batch.
__global__ void k_reset_mut(Ctl *c) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c->touched = 0;
    c->small_cnt = c->big_cnt = c->add_cnt = c->copy_cnt = c->overflow_cnt = 0;
    c->dtile_cnt = c->ht_cursor = 0;
    c->group_cursor = 0;
}

// InitGraph (driver.cuh:186-206): traversal state words and the first frontier.
// kind 0: single source (SSSP/BFS), kind 1: vertex seeds (CC, label = own id).
__global__ void k_init_traversal(uint64_t *state, uint32_t *stamp, uint32_t n, uint32_t kind,
                                 uint32_t source, uint32_t *q, Ctl *c) {
    const uint32_t r = c->round;
    for (uint32_t v = gthread(); v < n; v += gthreads()) {
        if (kind == 1) {
            state[v] = (uint64_t(v) << 32) | kNone;
            stamp[v] = r;
            q[v] = v;
        } else {
            state[v] = v == source ? (uint64_t(0) << 32 | kNone) : ~0ull;
            stamp[v] = v == source ? r : 0;
        }
    }
    if (gthread() == 0) {
        if (kind == 0) q[0] = source;
        c->qcount[r % 3] = kind == 1 ? n : 1;
    }
}

// pr::Initialize (pr_residual.cuh:44-53) with the Kahan compensation array.
__global__ void k_init_pr(float *rank, float *comp, float *res, uint32_t *queued, uint32_t *q,
                          uint32_t n, Ctl *c) {
    const uint32_t r = c->round;
    for (uint32_t v = gthread(); v < n; v += gthreads()) {
        rank[v] = 0.0f;
        comp[v] = 0.0f;
        res[v] = 0.15f;
        queued[v] = 1u;
        q[v] = v;
    }
    if (gthread() == 0) c->qcount[r % 3] = n;
}

// ------------------------------------------------------------------ mutation
struct MutArgs {
    Ctl *ctl;
    Row *rows;
    Pools pools;
    uint32_t n;
    const cig_edge *dels;
    uint32_t nd;
    const cig_edge *adds;
    uint32_t na;
    uint64_t *cnt64;   // per vertex: deletes (low) | adds (high) of this call
    uint64_t *fill64;  // per vertex scatter cursors, same packing
    uint64_t *gofs;    // per vertex offset of its group in `grouped`
    uint32_t *touched;
    uint32_t *grouped;
    uint32_t *small_list;
    uint32_t *big_list;
    uint32_t *add_list;
    uint32_t *overflow;
    // long-row deletes
    uint32_t *htoff;   // per vertex: offset of its delete table (kNone: none)
    uint32_t *hkeys;   // delete tables: destination + 1 (0 = empty)
    uint2 *hcnts;      // delete tables: (requested, claimed)
    uint32_t *hcnt;    // per vertex: holes found (zero between calls)
    uint2 *dtiles;     // (row, first slot) of each delete scan tile
    struct AddPiece *pieces;  // copy pieces of the insert phase (AddPiece below)
    StatShard *stats;
};

// Rows with more than 64 deletes, or longer than 4096 slots, are deleted
// from in tiles of kDelTile slots spread over the whole grid. Their delete
// multisets live in open-addressing tables of 2^bits >= 2*nd slots.
constexpr uint32_t kDelTile = 4096;
constexpr uint32_t kDelLdsKeys = 8192;

__device__ __forceinline__ bool del_is_small(uint32_t nd, uint32_t deg) { return nd <= 64 && deg <= 4096; }

__device__ __forceinline__ uint32_t del_table_bits(uint32_t nd) {
    uint32_t bits = 4;
    while ((1ull << bits) < 2ull * nd) ++bits;
    return bits;
}

__device__ __forceinline__ uint32_t ht_hash(uint32_t x, uint32_t bits) {
    return (x * 0x9E3779B1u) >> (32 - bits);
}

// GroupedUpdateBatch (effective_update_batch.h:35-87) on the device: count
// records per source and collect the touched sources (one list reservation
// per workgroup step).
__global__ void k_count(MutArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    const uint32_t total = a.nd + a.na;
    uint64_t bad_total = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        bool first = false;
        uint32_t u = 0;
        bool bad = false;
        if (i < total) {
            const bool del = i < a.nd;
            const cig_edge e = del ? a.dels[i] : a.adds[i - a.nd];
            if (e.u >= a.n || e.v >= a.n) {
                bad = true;
            } else {
                u = e.u;
                const uint64_t old = atomicAdd((unsigned long long *)&a.cnt64[u],
                                               del ? 1ull : (1ull << 32));
                first = old == 0;
            }
        }
        bad_total += uint32_t(__popcll(__ballot(bad)));
        block_append(first, u, a.touched, &a.ctl->touched, scratch);
    }
    stat_add(&stat_shard(a.stats)->invalid, bad_total);
}

// Reserves each touched source's group and classifies its work; long-row
// deletes also reserve their table and scan tiles. Every cursor and list
// takes one atomic per workgroup step.
__global__ void k_plan(MutArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    __shared__ unsigned long long scratch64[kMaxWavesPerBlock + 1];
    const uint32_t touched = ld_acquire(&a.ctl->touched);
    for (uint32_t base = blockIdx.x * blockDim.x; base < touched; base += gthreads()) {
        const uint32_t t = base + threadIdx.x;
        bool small = false, big = false, add = false;
        uint32_t u = 0, len = 0, nd = 0, na = 0, deg = 0;
        if (t < touched) {
            u = a.touched[t];
            const uint64_t c = a.cnt64[u];
            nd = uint32_t(c);
            na = uint32_t(c >> 32);
            len = nd + na;
            deg = a.rows[u].deg;
            // one wave scans a row of up to 4096 slots for up to 64 deletes;
            // anything larger is cut into grid-wide tiles
            small = nd != 0 && del_is_small(nd, deg);
            big = nd != 0 && !small;
            add = na != 0;
        }
        const uint64_t g = block_reserve<unsigned long long>(len, (unsigned long long *)&a.ctl->group_cursor, scratch64);
        if (t < touched) a.gofs[u] = g;
        const uint32_t tsz = big ? (1u << del_table_bits(nd)) : 0u;
        const uint32_t ntl = big ? (deg + kDelTile - 1) / kDelTile : 0u;
        const uint32_t hb = block_reserve<uint32_t>(tsz, &a.ctl->ht_cursor, scratch);
        const uint32_t tb = block_reserve<uint32_t>(ntl, &a.ctl->dtile_cnt, scratch);
        if (t < touched) a.htoff[u] = big ? hb : kNone;
        for (uint32_t j = 0; j < ntl; ++j) a.dtiles[tb + j] = make_uint2(u, j * kDelTile);
        block_append(small, u, a.small_list, &a.ctl->small_cnt, scratch);
        block_append(big, u, a.big_list, &a.ctl->big_cnt, scratch);
        block_append(add, u, a.add_list, &a.ctl->add_cnt, scratch);
    }
}

// Groups the records by source; deletes of long rows also enter their
// row's table (destination -> requested count).
__global__ void k_scatter(MutArgs a) {
    const uint32_t total = a.nd + a.na;
    for (uint32_t i = gthread(); i < total; i += gthreads()) {
        const bool del = i < a.nd;
        const cig_edge e = del ? a.dels[i] : a.adds[i - a.nd];
        if (e.u >= a.n || e.v >= a.n) continue;
        const uint64_t pos = atomicAdd((unsigned long long *)&a.fill64[e.u],
                                       del ? 1ull : (1ull << 32));
        const uint32_t nd = uint32_t(a.cnt64[e.u]);
        const uint32_t slot = del ? uint32_t(pos) : nd + uint32_t(pos >> 32);
        a.grouped[a.gofs[e.u] + slot] = e.v;
        const uint32_t to = del ? a.htoff[e.u] : kNone;
        if (to != kNone) {
            // at most nd distinct keys in 2^bits >= 2*nd slots: the probe ends
            const uint32_t bits = del_table_bits(nd), mask = (1u << bits) - 1u;
            uint32_t h = ht_hash(e.v, bits);
            for (;;) {
                const uint32_t prev = atomicCAS(&a.hkeys[to + h], 0u, e.v + 1u);
                if (prev == 0u || prev == e.v + 1u) break;
                h = (h + 1u) & mask;
            }
            atomicAdd(&a.hcnts[to + h].x, 1u);
        }
    }
}

// Deletes for a row of at most 4096 slots with at most 64 delete records,
// one wave per row. Lane j holds delete j; the row streams through 64*4 slots at a
// time and each pending delete claims the first live occurrence in row order
// (BuildDeletionPlan, source_local_chunk_store.h:670-723); the scan stops once
// every delete found its edge. The k holes are then refilled from the row's
// last k slots, so the row is rewritten in O(k) stores instead of the
// reference's order-preserving compaction (CompactSourceInPlace, :733-780):
// the multiset of survivors is the same, their order may differ.
__global__ __launch_bounds__(kBlock) void k_del_rows(MutArgs a) {
    __shared__ uint32_t slot_lds[kWavesPerBlock][kWave];
    uint32_t *slots = slot_lds[wave_in_block()];
    const uint32_t cnt = ld_acquire(&a.ctl->small_cnt);
    const uint32_t lane = lane_id();
    uint64_t acc_applied = 0, acc_missing = 0;
    constexpr int U = 4;
    for (uint32_t j = gwave(); j < cnt; j += gwaves()) {
        const uint32_t u = a.small_list[j];
        const Row r = a.rows[u];
        const uint32_t nd = uint32_t(a.cnt64[u]);
        const uint32_t *dl = a.grouped + a.gofs[u];
        const uint32_t mine = lane < nd ? dl[lane] : kNone;
        const uint64_t all = nd >= 64 ? ~0ull : ((1ull << nd) - 1);
        uint64_t pend = all;
        uint32_t hole = kNone;
        uint32_t *row = a.pools.row(r.off);
        for (uint32_t base = 0; base < r.deg && pend; base += kWave * U) {
            uint32_t e[U];
            bool live[U];
#pragma unroll
            for (int q = 0; q < U; ++q) {
                const uint32_t p = base + uint32_t(q) * kWave + lane;
                live[q] = p < r.deg;
                e[q] = live[q] ? row[p] : kNone;
            }
            for (uint64_t todo = pend; todo; todo &= todo - 1) {
                const int jj = __ffsll((unsigned long long)todo) - 1;
                const uint32_t d = __shfl(mine, jj, kWave);
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    const uint64_t m = __ballot(live[q] && e[q] == d);
                    if (m) {
                        const uint32_t l = uint32_t(__ffsll((unsigned long long)m) - 1);
                        if (lane == l) live[q] = false;
                        if (lane == uint32_t(jj)) hole = base + uint32_t(q) * kWave + l;
                        pend &= ~(1ull << jj);
                        break;
                    }
                }
            }
        }
        const uint64_t matched = all & ~pend;
        const uint32_t k = uint32_t(__popcll(matched));
        if (k) {
            const uint32_t newdeg = r.deg - k;
            // tail slot newdeg+lane (lane < k): dead if it is one of the holes
            bool tail_dead = false;
            for (uint64_t mm = matched; mm; mm &= mm - 1) {
                const uint32_t h = __shfl(hole, __ffsll((unsigned long long)mm) - 1, kWave);
                tail_dead |= (lane < k) && h == newdeg + lane;
            }
            const bool tail_live = lane < k && !tail_dead;
            const uint32_t tval = tail_live ? row[newdeg + lane] : 0u;
            const bool front = ((matched >> lane) & 1ull) && hole < newdeg;
            const uint64_t fm = __ballot(front);
            const uint64_t lm = __ballot(tail_live);
            if (front) slots[__popcll(fm & __lanemask_lt())] = hole;
            wave_lds_sync();
            if (tail_live) row[slots[__popcll(lm & __lanemask_lt())]] = tval;
            wave_lds_sync();
            if (lane == 0) a.rows[u].deg = newdeg;
        }
        acc_applied += k;
        acc_missing += nd - k;
    }
    StatShard *sh = stat_shard(a.stats);
    stat_add(&sh->deletes_applied, acc_applied);
    stat_add(&sh->missing, acc_missing);
}

// Exclusive scan of a 0/1 flag over a 256-thread workgroup; *total gets the
// sum. Contains barriers: every thread of the workgroup must call it.
__device__ __forceinline__ uint32_t block_flag_scan(bool f, uint32_t *wsum, uint32_t &total) {
    const uint64_t m = __ballot(f);
    const uint32_t w = wave_in_block();
    if (lane_id() == 0) wsum[w] = uint32_t(__popcll(m));
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (uint32_t i = 0; i < blockDim.x / kWave; ++i) {
        if (i < w) before += wsum[i];
        total += wsum[i];
    }
    __syncthreads();
    return before + uint32_t(__popcll(m & __lanemask_lt()));
}

// Long-row deletes, scan phase: one workgroup per kDelTile-slot tile. A slot
// whose destination is in the row's table claims one requested occurrence
// (first claims win; equal destinations are interchangeable, so the survivor
// multiset is the reference's, BuildDeletionPlan,
// source_local_chunk_store.h:670-723); a claimed slot is overwritten with
// kNone and its position is recorded in the row's group (whose delete
// records are no longer needed once the table is built).
__global__ __launch_bounds__(256) void k_del_scan(MutArgs a) {
    __shared__ uint32_t keys[kDelLdsKeys];
    const uint32_t nt = ld_acquire(&a.ctl->dtile_cnt);
    const uint32_t tid = threadIdx.x;
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        const uint2 dt = a.dtiles[t];
        const uint32_t u = dt.x, start = dt.y;
        const Row r = a.rows[u];
        const uint32_t nd = uint32_t(a.cnt64[u]);
        const uint32_t bits = del_table_bits(nd), mask = (1u << bits) - 1u;
        const uint32_t to = a.htoff[u];
        const bool in_lds = (1u << bits) <= kDelLdsKeys;
        __syncthreads();  // the previous tile is done with `keys`
        if (in_lds)
            for (uint32_t k = tid; k <= mask; k += blockDim.x) keys[k] = a.hkeys[to + k];
        __syncthreads();
        const uint32_t *kt = in_lds ? keys : a.hkeys + to;
        uint32_t *row = a.pools.row(r.off) + start;  // start is a multiple of 4
        const uint32_t len = min(kDelTile, r.deg - start);
        const uint32_t nq = (len + 3u) >> 2;
        uint32_t *holes = a.grouped + a.gofs[u];
        for (uint32_t q0 = 0; q0 < nq; q0 += blockDim.x * 4) {
            uint32_t d[4][4], live[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t q = q0 + uint32_t(i) * blockDim.x + tid;
                live[i] = q < nq ? load_quad(row, q, len, d[i]) : 0u;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t q = q0 + uint32_t(i) * blockDim.x + tid;
                for (uint32_t sl = 0; sl < live[i]; ++sl) {
                    const uint32_t x = d[i][sl];
                    uint32_t h = ht_hash(x, bits);
                    for (uint32_t probe = 0; probe <= mask; ++probe) {
                        const uint32_t key = kt[h];
                        if (key == 0u) break;
                        if (key == x + 1u) {
                            if (atomicAdd(&a.hcnts[to + h].y, 1u) < a.hcnts[to + h].x) {
                                row[4u * q + sl] = kNone;
                                holes[atomicAdd(&a.hcnt[u], 1u)] = start + 4u * q + sl;
                            }
                            break;
                        }
                        h = (h + 1u) & mask;
                    }
                }
            }
        }
    }
}

// Long-row deletes, fix-up phase: one workgroup per row. With k holes the
// row shrinks to deg-k; the live slots of the last k positions move into the
// holes below deg-k (O(k) stores; survivors keep their multiset, not their
// order, unlike CompactSourceInPlace, source_local_chunk_store.h:733-780).
// The row's table and hole counter are cleared for the next call.
__global__ __launch_bounds__(256) void k_del_fix(MutArgs a) {
    __shared__ uint32_t wsum[256 / kWave];
    const uint32_t cnt = ld_acquire(&a.ctl->big_cnt);
    const uint32_t tid = threadIdx.x;
    uint64_t applied = 0, missing = 0;
    for (uint32_t j = blockIdx.x; j < cnt; j += gridDim.x) {
        const uint32_t u = a.big_list[j];
        const Row r = a.rows[u];
        const uint32_t nd = uint32_t(a.cnt64[u]);
        const uint32_t k = a.hcnt[u];
        uint32_t *hl = a.grouped + a.gofs[u];
        uint32_t *row = a.pools.row(r.off);
        const uint32_t newdeg = r.deg - k;
        // 1. compact the holes below newdeg to the front of the hole list
        //    (write index <= read index, and the scan's barrier orders them)
        uint32_t run = 0;
        for (uint32_t b = 0; b < k; b += blockDim.x) {
            const uint32_t i = b + tid;
            const uint32_t h = i < k ? hl[i] : 0u;
            const bool front = i < k && h < newdeg;
            uint32_t tot;
            const uint32_t rk = block_flag_scan(front, wsum, tot);
            if (front) hl[run + rk] = h;
            run += tot;
        }
        __syncthreads();
        // 2. live slots of [newdeg, deg) fill those holes in order
        run = 0;
        for (uint32_t b = 0; b < k; b += blockDim.x) {
            const uint32_t i = b + tid;
            const uint32_t v = i < k ? row[newdeg + i] : kNone;
            const bool live = v != kNone;
            uint32_t tot;
            const uint32_t rk = block_flag_scan(live, wsum, tot);
            if (live) row[hl[run + rk]] = v;
            run += tot;
        }
        // 3. clear the table for the next call
        const uint32_t to = a.htoff[u], tsz = 1u << del_table_bits(nd);
        for (uint32_t s = tid; s < tsz; s += blockDim.x) {
            a.hkeys[to + s] = 0u;
            a.hcnts[to + s] = make_uint2(0u, 0u);
        }
        if (tid == 0) {
            a.rows[u].deg = newdeg;
            a.hcnt[u] = 0u;
        }
        applied += k;
        missing += nd - k;
        __syncthreads();
    }
    if (wave_in_block() == 0) {
        StatShard *sh = stat_shard(a.stats);
        stat_add(&sh->deletes_applied, applied);
        stat_add(&sh->missing, missing);
    }
}

// Insert phase (PatchOrInvalidateCachedAdjacency's grow path,
// csr_graph.cuh:1793-1821, and RewriteSourceToNewBlock,
// source_local_chunk_store.h:782-848): each touched source appends its adds;
// a row that outgrows its block moves to a block of the next power of two at
// the tail of its tier. k_add_plan decides per row (one thread each),
// reserves pool blocks and copy pieces with one atomic per workgroup and
// tier, and rewrites the descriptors; k_add_copy then moves every piece with
// whole waves. A row that cannot be placed is listed in `overflow` and left
// for the host to retry after compaction.
struct AddPiece {
    const uint32_t *src;
    uint32_t *dst;
    uint32_t len;
    uint32_t pad;
};
constexpr uint32_t kPieceLen = 1024;

__device__ __forceinline__ uint32_t pieces_of(uint64_t len) { return uint32_t((len + kPieceLen - 1) / kPieceLen); }

__global__ void k_add_plan(MutArgs a, const uint32_t *list, const uint32_t *list_count) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    __shared__ unsigned long long scratch64[kMaxWavesPerBlock + 1];
    const uint32_t cnt = ld_acquire(list_count);
    uint64_t acc_adds = 0, acc_reloc = 0, acc_garbage[2] = {0, 0};
    for (uint32_t base = blockIdx.x * blockDim.x; base < cnt; base += gthreads()) {
        const uint32_t j = base + threadIdx.x;
        const bool on = j < cnt;
        uint32_t u = 0, nd = 0, na = 0, tier = 0;
        Row r{};
        uint64_t newdeg = 0, newcap = 0;
        bool grow = false;
        if (on) {
            u = list[j];
            const uint64_t c = a.cnt64[u];
            nd = uint32_t(c);
            na = uint32_t(c >> 32);
            r = a.rows[u];
            newdeg = uint64_t(r.deg) + na;
            tier = (r.off & kHostTier) ? 1u : 0u;
            grow = newdeg > r.cap;
            if (grow) newcap = next_pow2_dev(newdeg);
        }
        const uint64_t o0 = block_reserve<unsigned long long>(grow && tier == 0 ? newcap : 0ull,
                                                              (unsigned long long *)&a.ctl->pool_tail[0], scratch64);
        const uint64_t o1 = block_reserve<unsigned long long>(grow && tier == 1 ? newcap : 0ull,
                                                              (unsigned long long *)&a.ctl->pool_tail[1], scratch64);
        const uint64_t o = tier ? o1 : o0;
        const bool over = grow && (o + newcap > a.ctl->pool_cap[tier] || newcap > 0xFFFFFFFFull);
        block_append(over, u, a.overflow, &a.ctl->overflow_cnt, scratch);
        const bool go = on && !over && na != 0;
        const uint64_t noff = grow ? (o | (tier ? kHostTier : 0ull)) : r.off;
        const uint32_t np = go ? (grow ? pieces_of(r.deg) : 0u) + pieces_of(na) : 0u;
        const uint32_t pb = block_reserve<uint32_t>(np, &a.ctl->copy_cnt, scratch);
        if (go) {
            const uint32_t *old_row = a.pools.row(r.off);
            uint32_t *row = a.pools.row(noff);
            uint32_t k = pb;
            if (grow)
                for (uint32_t p = 0; p < r.deg; p += kPieceLen)
                    a.pieces[k++] = AddPiece{old_row + p, row + p, min(kPieceLen, r.deg - p), 0u};
            const uint32_t *al = a.grouped + a.gofs[u] + nd;
            for (uint32_t p = 0; p < na; p += kPieceLen)
                a.pieces[k++] = AddPiece{al + p, row + r.deg + p, min(kPieceLen, na - p), 0u};
            if (grow) {
                Row nr;
                nr.off = noff;
                nr.deg = uint32_t(newdeg);
                nr.cap = uint32_t(newcap);
                a.rows[u] = nr;
                acc_garbage[tier] += r.cap;
                acc_reloc += 1;
            } else {
                a.rows[u].deg = uint32_t(newdeg);
            }
            acc_adds += na;
        }
    }
    StatShard *sh = stat_shard(a.stats);
    stat_add(&sh->adds_applied, wave_sum64(acc_adds));
    stat_add(&sh->relocated, wave_sum64(acc_reloc));
    stat_add(&sh->garbage[0], wave_sum64(acc_garbage[0]));
    stat_add(&sh->garbage[1], wave_sum64(acc_garbage[1]));
}

// Moves the pieces of the insert phase: a wave takes 64 pieces, scans their
// lengths and copies the concatenation 256 slots per step, so short pieces
// (a few adds) and long ones (a relocated row) keep every lane busy.
__global__ __launch_bounds__(kBlock) void k_add_copy(MutArgs a) {
    __shared__ uint32_t pref_lds[kWavesPerBlock][kWave];
    uint32_t *pref = pref_lds[wave_in_block()];
    const uint32_t cnt = ld_acquire(&a.ctl->copy_cnt);
    const uint32_t lane = lane_id();
    for (uint64_t base = uint64_t(gwave()) * kWave; base < cnt; base += uint64_t(gwaves()) * kWave) {
        const uint32_t i = uint32_t(base) + lane;
        AddPiece pc{nullptr, nullptr, 0u, 0u};
        if (i < cnt) pc = a.pieces[i];
        const uint32_t incl = wave_incl_scan(pc.len);
        const uint32_t total = __shfl(incl, kWave - 1, kWave);
        pref[lane] = incl - pc.len;
        wave_lds_sync();
        const unsigned long long src = (unsigned long long)pc.src, dst = (unsigned long long)pc.dst;
        for (uint32_t k0 = 0; k0 < total; k0 += kWave * 4) {
            uint32_t val[4], *dp[4];
            bool ok[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t k = k0 + uint32_t(e) * kWave + lane;
                ok[e] = k < total;
                const uint32_t kk = ok[e] ? k : total - 1;
                const uint32_t own = owner_of(pref, kk);
                const uint32_t off = kk - pref[own];
                const uint32_t *sp = reinterpret_cast<const uint32_t *>(__shfl(src, int(own), kWave));
                dp[e] = reinterpret_cast<uint32_t *>(__shfl(dst, int(own), kWave)) + off;
                val[e] = sp[off];
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (ok[e]) *dp[e] = val[e];
        }
        wave_lds_sync();
    }
}

__global__ void k_clear_touched(MutArgs a) {
    const uint32_t touched = ld_acquire(&a.ctl->touched);
    for (uint32_t t = gthread(); t < touched; t += gthreads()) {
        const uint32_t u = a.touched[t];
        a.cnt64[u] = 0;
        a.fill64[u] = 0;
    }
}

// ---------------------------------------------------- deletion invalidation
struct InvalArgs {
    Ctl *ctl;
    const Row *rows;
    uint64_t *state;
    uint32_t *stamp;
    uint32_t *aff;
    uint32_t *alist;
    uint32_t *filt;  // global copy of the repair filter (kFilterWords words)
    uint32_t *q0, *q1;
    Tile *t0, *t1;
    ExpandCfg cfg;
    const cig_edge *dels;
    uint32_t nd;
    const cig_edge *adds;
    uint32_t na;
    uint32_t n;
    uint32_t kind;  // 0 = single source (SSSP/BFS), 1 = vertex seeds (CC)
    uint32_t max_rounds;
    uint64_t seed;
};

// reset_del_edges (driver.cuh:371-411): a deleted (u,v) whose v has parent u
// seeds the affected set. Parallel copies of (u,v) are treated conservatively.
__global__ void k_inval_seed(InvalArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.nd; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        bool hit = false;
        uint32_t v = 0;
        if (i < a.nd) {
            const cig_edge e = a.dels[i];
            if (e.u < a.n && e.v < a.n && uint32_t(a.state[e.v]) == e.u) {
                v = e.v;
                hit = bit_claim(a.aff, v);
            }
        }
        block_append(hit, v, a.alist, &a.ctl->acount, scratch);
    }
}

// Invalidation waves (RunSyncPushDDB_del rounds, framework.cuh:2815-2853) as
// one persistent launch: the affected list grows in place; round k expands
// the items appended by round k-1.
template <int QPL>
__global__ __launch_bounds__(kBlock) void k_inval_loop(InvalArgs a) {
    __shared__ WaveLds lds[kWavesPerBlock];
    WaveLds &L = lds[wave_in_block()];
    Ctl *c = a.ctl;
    InvalidateOp op{a.rows, a.state, a.aff};
    Counters ctr;
    uint32_t b = 0;
    for (uint32_t k = 0; k < a.max_rounds; ++k) {
        const uint32_t e = ld_acquire(&c->acount);
        const uint32_t nt = ld_acquire(&c->icount[k % 3]);
        if (b == e && nt == 0) break;
        if (blockIdx.x == 0 && threadIdx.x == 0) c->icount[(k + 2) % 3] = 0;
        const DirectOut out{Sink{a.alist, &c->acount}, TileSink{(k & 1) ? a.t0 : a.t1, &c->icount[(k + 1) % 3]}};
        expand_range<QPL>(op, a.cfg, a.alist + b, e - b, out, L, ctr);
        expand_tiles<QPL>(op, a.cfg, (k & 1) ? a.t1 : a.t0, nt, out);
        b = e;
        cg::this_grid().sync();
    }
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
}

// Opens the invalidation frontier after the seeds: [0, acount).
__global__ void k_inval_begin(Ctl *c) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c->ibegin = 0;
    c->iend = c->acount;
    c->iround = 0;
    c->icount[0] = c->icount[1] = c->icount[2] = 0;
}

// One invalidation wave as an ordinary launch: expands [ibegin, iend) of the
// affected list; the closing workgroup moves the window to the items this
// round appended.
using RoundStage = StageLds<2048, 128>;

template <int QPL>
__global__ __launch_bounds__(kBlock) void k_inval_round(InvalArgs a, uint64_t *prog, uint64_t tag) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ RoundStage stage;
    Ctl *c = a.ctl;
    const uint32_t b = c->ibegin, e = c->iend, k = c->iround;
    const uint32_t nt = c->icount[k % 3];
    if (b == e && nt == 0) return post_idle(prog, tag);
    InvalidateOp op{a.rows, a.state, a.aff};
    Counters ctr;
    const StagedOut<2048, 128> out{Sink{a.alist, &c->acount},
                                   TileSink{(k & 1) ? a.t0 : a.t1, &c->icount[(k + 1) % 3]}, &stage};
    out_init(out);
    expand_range<QPL>(op, a.cfg, a.alist + b, e - b, out, lds[wave_in_block()], ctr);
    expand_tiles<QPL>(op, a.cfg, (k & 1) ? a.t1 : a.t0, nt, out);
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
    close_round(c, [&] {
        c->icount[k % 3] = 0;
        c->ibegin = e;
        c->iend = ld_acquire(&c->acount);
        c->iround = k + 1;
        post_progress(prog, tag, c->iend != e || c->icount[(k + 1) % 3] != 0);
    });
}

// reset_affected_values (driver.cuh:418-433): affected vertices restart from
// the initial value (CC: own label), and all of them form the next frontier.
__global__ void k_reset_affected(InvalArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    const uint32_t cnt = ld_acquire(&a.ctl->acount);
    const uint32_t r = a.ctl->round;
    uint32_t *q = (r & 1) ? a.q1 : a.q0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < cnt; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        bool push = false;
        uint32_t v = 0;
        if (i < cnt) {
            v = a.alist[i];
            a.state[v] = a.kind == 1 ? ((uint64_t(v) << 32) | kNone) : ~0ull;
            push = claim(a.stamp, v, r);
        }
        block_append(push, v, q, &a.ctl->qcount[r % 3], scratch);
    }
}

__global__ void k_clear_affected(InvalArgs a) {
    const uint32_t cnt = ld_acquire(&a.ctl->acount);
    for (uint32_t i = gthread(); i < cnt; i += gthreads()) a.aff[a.alist[i] >> 5] = 0u;
}

// SeedAddedEdgeFrontier (framework.cuh:238-294): each inserted (u,v) offers
// value(u)+w to v; improved destinations join the pending frontier.
template <int W, bool kStrict>
__global__ void k_insert_seed(InvalArgs a, uint32_t skip_affected) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    const uint32_t r = a.ctl->round;
    uint32_t *q = (r & 1) ? a.q1 : a.q0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.na; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        bool push = false;
        uint32_t v = 0;
        if (i < a.na) {
            const cig_edge e = a.adds[i];
            if (e.u < a.n && e.v < a.n && !(skip_affected && bit_test(a.aff, e.u))) {
                const uint32_t val = uint32_t(a.state[e.u] >> 32);
                if (val != kInf) {
                    const uint64_t cand = uint64_t(val) + weight<W>(e.u, e.v, a.seed);
                    if (cand < kInf && relax<kStrict>(a.state, e.v, uint32_t(cand), e.u)) {
                        v = e.v;
                        push = claim(a.stamp, v, r);
                    }
                }
            }
        }
        block_append(push, v, q, &a.ctl->qcount[r % 3], scratch);
    }
}

// -------------------------------------------------------------- repair scan
// The affected list is folded once into a bitmap filter of 2^18 bits (a
// multiplicative hash of the id) in global memory; every repair workgroup
// copies it into LDS. An edge whose destination misses the filter is not
// affected, so almost every edge of the scan is settled in LDS and only
// filter hits read the global affected bitmap. The filter is off when the
// affected list is too long to stay selective.
constexpr uint32_t kFilterBits = 18;
constexpr uint32_t kFilterWords = (1u << kFilterBits) / 32;
constexpr uint32_t kFilterMaxItems = 1u << 16;
constexpr uint32_t kRepairBlock = 512;

__device__ __forceinline__ uint32_t filter_hash(uint32_t v) { return (v * 0x9E3779B1u) >> (32 - kFilterBits); }

struct AffFilter {
    const uint32_t *lds;  // nullptr: filter off
    const uint32_t *aff;
    __device__ __forceinline__ bool test(uint32_t v) const {
        if (lds) {
            const uint32_t h = filter_hash(v);
            if (!((lds[h >> 5] >> (h & 31)) & 1u)) return false;
        }
        return bit_test(aff, v);
    }
};

// a.filt must be zero on entry (k_fill32 before this launch).
__global__ void k_build_filter(InvalArgs a) {
    const uint32_t na = ld_acquire(&a.ctl->acount);
    if (na > kFilterMaxItems) return;
    for (uint32_t i = gthread(); i < na; i += gthreads()) {
        const uint32_t h = filter_hash(a.alist[i]);
        atomicOr(&a.filt[h >> 5], 1u << (h & 31));
    }
}

__device__ __forceinline__ bool load_filter(uint32_t *lds, const uint32_t *g, uint32_t na) {
    if (na > kFilterMaxItems) return false;
    const uint4 *src = reinterpret_cast<const uint4 *>(g);
    uint4 *dst = reinterpret_cast<uint4 *>(lds);
    for (uint32_t w = threadIdx.x; w < kFilterWords / 4; w += blockDim.x) dst[w] = src[w];
    __syncthreads();
    return true;
}

template <int W, bool kStrict>
struct RepairScanOp {
    const Row *rows;
    uint64_t *state;
    AffFilter filt;
    uint64_t seed;
    __device__ bool source(uint32_t v, uint32_t, Row &row, uint32_t &pay) const {
        if (filt.test(v)) return false;
        const uint32_t val = uint32_t(state[v] >> 32);
        if (val == kInf) return false;
        row = rows[v];
        pay = val;
        return row.deg != 0;
    }
    __device__ bool tile_source(const Tile &t, uint32_t &pay) const {
        pay = uint32_t(state[t.v] >> 32);
        return pay != kInf;
    }
    // the LDS filter answers most edges without a global access
    __device__ uint64_t peek(uint32_t y) const {
        if (!filt.lds) return 1;
        const uint32_t h = filter_hash(y);
        return (filt.lds[h >> 5] >> (h & 31)) & 1u;
    }
    __device__ bool edge(uint32_t u, uint32_t val, uint32_t y, uint64_t maybe) const {
        if (!maybe || !bit_test(filt.aff, y)) return false;
        const uint64_t c = uint64_t(val) + weight<W>(u, y, seed);
        if (c < kInf) relax<kStrict>(state, y, uint32_t(c), u);
        return false;
    }
};

template <int W, bool kStrict, int QPL>
__global__ __launch_bounds__(kRepairBlock) void k_repair_scan(InvalArgs a) {
    __shared__ __align__(16) uint32_t filt[kFilterWords];
    __shared__ WaveLds lds[kRepairBlock / kWave];
    __shared__ StageLds<2, 160> stage;
    Ctl *c = a.ctl;
    const uint32_t na = ld_acquire(&c->acount);
    if (na == 0) return;
    const bool on = load_filter(filt, a.filt, na);
    RepairScanOp<W, kStrict> op{a.rows, a.state, AffFilter{on ? filt : nullptr, a.aff}, a.seed};
    Counters ctr;
    const StagedOut<2, 160> out{Sink{nullptr, &c->scount}, TileSink{a.t0, &c->scount}, &stage};
    out_init(out);
    expand_range<QPL>(op, a.cfg, nullptr, a.n, out, lds[wave_in_block()], ctr);
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
}

template <int W, bool kStrict, int QPL>
__global__ __launch_bounds__(kRepairBlock) void k_repair_tiles(InvalArgs a) {
    __shared__ __align__(16) uint32_t filt[kFilterWords];
    Ctl *c = a.ctl;
    const uint32_t na = ld_acquire(&c->acount);
    const uint32_t nt = ld_acquire(&c->scount);
    if (na == 0 || nt == 0) return;
    const bool on = load_filter(filt, a.filt, na);
    RepairScanOp<W, kStrict> op{a.rows, a.state, AffFilter{on ? filt : nullptr, a.aff}, a.seed};
    expand_tiles<QPL>(op, a.cfg, a.t0, nt, DirectOut{Sink{nullptr, &c->scount}, TileSink{nullptr, &c->scount}});
}

// ------------------------------------------------- one-shot expansion passes
template <int QPL, class Op>
__global__ __launch_bounds__(kBlock) void k_expand_once(Op op, ExpandCfg cfg, const uint32_t *items,
                                                         const uint32_t *count_ptr, uint32_t count,
                                                         Sink sink, TileSink ts, const uint32_t *gate) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ RoundStage stage;
    if (gate && ld_acquire(gate) == 0) return;
    Counters ctr;
    const uint32_t n = count_ptr ? ld_acquire(count_ptr) : count;
    const StagedOut<2048, 128> out{sink, ts, &stage};
    out_init(out);
    expand_range<QPL>(op, cfg, items, n, out, lds[wave_in_block()], ctr);
    flush_counters(ctr, cfg.stats, cfg.stage);
    out_flush(out);
}

template <int QPL, class Op>
__global__ __launch_bounds__(kBlock) void k_expand_tiles(Op op, ExpandCfg cfg, const Tile *tiles,
                                                          const uint32_t *ntiles, Sink sink) {
    __shared__ RoundStage stage;
    const StagedOut<2048, 128> out{sink, TileSink{nullptr, sink.count}, &stage};
    out_init(out);
    expand_tiles<QPL>(op, cfg, tiles, ld_acquire(ntiles), out);
    out_flush(out);
}

// ------------------------------------------------------ traversal fixed point
struct LoopArgs {
    Ctl *ctl;
    const Row *rows;
    uint64_t *state;
    uint32_t *stamp;
    uint32_t *hot;
    uint32_t *q0, *q1;
    Tile *t0, *t1;
    ExpandCfg cfg;
    uint64_t seed;
    uint32_t max_rounds;
};

// One persistent cooperative launch runs rounds until the frontier and the
// deferred tiles are empty (RunExactGpuClosure's grid.sync loop,
// framework.cuh:342-454, without host round trips). Round r consumes
// queue[r&1] / count slot r%3 and produces into the other queue and slot
// (r+1)%3; slot (r+2)%3 is cleared for round r+1. Tiles of long rows found in
// round r are expanded in round r+1 (re-reading the row owner's value).
template <int W, bool kStrict, int QPL>
__global__ __launch_bounds__(kBlock) void k_relax_loop(LoopArgs a) {
    __shared__ WaveLds lds[kWavesPerBlock];
    WaveLds &L = lds[wave_in_block()];
    Ctl *c = a.ctl;
    uint32_t r = ld_acquire(&c->round);
    Counters ctr;
    uint32_t it = 0;
    for (; it < a.max_rounds; ++it, ++r) {
        const uint32_t nq = ld_acquire(&c->qcount[r % 3]);
        const uint32_t nt = ld_acquire(&c->tcount[r % 3]);
        if (nq == 0 && nt == 0) break;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            c->qcount[(r + 2) % 3] = 0;
            c->tcount[(r + 2) % 3] = 0;
        }
        RelaxOp<W, kStrict> op{a.rows, a.state, a.stamp, a.hot, a.seed, r + 1};
        const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
        const DirectOut out{Sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]},
                            TileSink{(r & 1) ? a.t0 : a.t1, &c->tcount[(r + 1) % 3]}};
        expand_range<QPL>(op, a.cfg, qin, nq, out, L, ctr);
        expand_tiles<QPL>(op, a.cfg, (r & 1) ? a.t1 : a.t0, nt, out);
        cg::this_grid().sync();
    }
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->round = r;
        c->rounds += it;
    }
}

// The last workgroup of a round kernel to finish runs `close` on thread 0.
// Every workgroup publishes its writes (fence) before counting itself done,
// so `close` sees all appends of the round; the next kernel in the stream
// sees what `close` wrote. No workgroup ever waits for another.
template <class F>
__device__ __forceinline__ void close_round(Ctl *c, F close) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&c->done, 1u) == gridDim.x - 1) {
            __threadfence();
            c->done = 0;
            close();
        }
    }
}

// One relax round as an ordinary launch (the default driver): the host
// keeps a few rounds enqueued ahead; a round whose frontier and deferred
// tiles are empty returns at once.
template <int W, bool kStrict, int QPL>
__global__ __launch_bounds__(kBlock) void k_relax_round(LoopArgs a, uint64_t *prog, uint64_t tag) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ RoundStage stage;
    Ctl *c = a.ctl;
    const uint32_t r = c->round;
    const uint32_t nq = c->qcount[r % 3], nt = c->tcount[r % 3];
    if (nq == 0 && nt == 0) return post_idle(prog, tag);
    Counters ctr;
    RelaxOp<W, kStrict> op{a.rows, a.state, a.stamp, a.hot, a.seed, r + 1};
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    const StagedOut<2048, 128> out{Sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]},
                                   TileSink{(r & 1) ? a.t0 : a.t1, &c->tcount[(r + 1) % 3]}, &stage};
    out_init(out);
    expand_range<QPL>(op, a.cfg, qin, nq, out, lds[wave_in_block()], ctr);
    expand_tiles<QPL>(op, a.cfg, (r & 1) ? a.t1 : a.t0, nt, out);
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
    close_round(c, [&] {
        c->qcount[r % 3] = 0;
        c->tcount[r % 3] = 0;
        c->round = r + 1;
        c->rounds += 1;
        post_progress(prog, tag, c->qcount[(r + 1) % 3] != 0 || c->tcount[(r + 1) % 3] != 0);
    });
}

// ---------------------------------------------------------- PageRank rounds
struct PrArgs {
    Ctl *ctl;
    const Row *rows;
    float *rank, *comp, *res, *delta;
    uint32_t *queued;
    uint32_t *q0, *q1;
    Tile *t0, *t1;
    ExpandCfg cfg;
    float eps;
    uint32_t max_rounds;
};

// Runtime::Converge (pr_residual.cuh:171-187) as one persistent launch:
// consume (rank += r with Kahan compensation, r = 0, contribution
// alpha*r/deg), grid barrier, scatter, grid barrier.
template <int QPL>
__global__ __launch_bounds__(kBlock) void k_pr_loop(PrArgs a) {
    __shared__ WaveLds lds[kWavesPerBlock];
    WaveLds &L = lds[wave_in_block()];
    Ctl *c = a.ctl;
    uint32_t r = ld_acquire(&c->round);
    Counters ctr;
    uint32_t it = 0;
    for (; it < a.max_rounds; ++it, ++r) {
        const uint32_t nq = ld_acquire(&c->qcount[r % 3]);
        if (nq == 0) break;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            c->qcount[(r + 2) % 3] = 0;
            c->tcount[(r + 2) % 3] = 0;
        }
        const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
        Tile *tiles = (r & 1) ? a.t1 : a.t0;
        for (uint32_t i = gthread(); i < nq; i += gthreads()) {
            const uint32_t v = qin[i];
            const float rv = a.res[v];
            a.queued[v] = 0u;
            float d = 0.0f;
            if (!isfinite(rv) || !isfinite(a.rank[v])) atomicExch(&c->pr_invalid, 1u);
            if (fabsf(rv) > a.eps) {
                a.res[v] = 0.0f;
                const float x = a.rank[v];
                const float y = rv - a.comp[v];
                const float t = x + y;
                a.comp[v] = (t - x) - y;
                a.rank[v] = t;
                const Row row = a.rows[v];
                ctr.vexp += 1;
                if (row.deg) {
                    d = 0.85f * rv / float(row.deg);
                    if (row.deg >= a.cfg.split) {
                        const uint32_t nt = (row.deg + a.cfg.tile_len - 1) / a.cfg.tile_len;
                        const uint32_t t0 = atomicAdd(&c->tcount[r % 3], nt);
                        for (uint32_t j = 0; j < nt; ++j) {
                            Tile t;
                            t.start = row.off + uint64_t(j) * a.cfg.tile_len;
                            t.v = v;
                            t.len = min(a.cfg.tile_len, row.deg - j * a.cfg.tile_len);
                            t.payload = __float_as_uint(d);
                            t.pad = 0;
                            tiles[t0 + j] = t;
                        }
                        ctr.scanned += row.deg;
                        if (row.off & kHostTier) ctr.host += row.deg;
                        d = 0.0f;
                    }
                }
            }
            a.delta[i] = d;
        }
        cg::this_grid().sync();
        PrScatterOp op{a.rows, a.res, a.queued, a.delta, &c->pr_invalid, a.eps};
        // tiles of long rows were made while consuming; the scatter makes none
        const DirectOut out{Sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]}, TileSink{tiles, &c->scount}};
        Counters sctr;
        expand_range<QPL>(op, a.cfg, qin, nq, out, L, sctr);
        ctr.scanned += sctr.scanned;
        ctr.host += sctr.host;
        expand_tiles<QPL>(op, a.cfg, tiles, ld_acquire(&c->tcount[r % 3]), out);
        cg::this_grid().sync();
    }
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->round = r;
        c->rounds += it;
        c->converged = ld_acquire(&c->qcount[r % 3]) == 0 ? 1u : 0u;
    }
}

// PageRank round, consume phase (pr_residual.cuh:54-71) as an ordinary
// launch; long rows are cut into tiles for this round's scatter (one tile
// reservation per workgroup step).
__global__ __launch_bounds__(kBlock) void k_pr_consume(PrArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    Ctl *c = a.ctl;
    const uint32_t r = c->round;
    const uint32_t nq = c->qcount[r % 3];
    if (nq == 0) return;
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    Tile *tiles = (r & 1) ? a.t1 : a.t0;
    Counters ctr;
    for (uint32_t base = blockIdx.x * blockDim.x; base < nq; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        uint32_t v = 0, nt = 0;
        Row row{};
        float d = 0.0f;
        if (i < nq) {
            v = qin[i];
            const float rv = a.res[v];
            a.queued[v] = 0u;
            if (!isfinite(rv) || !isfinite(a.rank[v])) atomicExch(&c->pr_invalid, 1u);
            if (fabsf(rv) > a.eps) {
                a.res[v] = 0.0f;
                const float x = a.rank[v];
                const float y = rv - a.comp[v];
                const float t = x + y;
                a.comp[v] = (t - x) - y;
                a.rank[v] = t;
                row = a.rows[v];
                ctr.vexp += 1;
                if (row.deg) {
                    d = 0.85f * rv / float(row.deg);
                    if (row.deg >= a.cfg.split) nt = (row.deg + a.cfg.tile_len - 1) / a.cfg.tile_len;
                }
            }
        }
        const uint32_t t0 = block_reserve<uint32_t>(nt, &c->tcount[r % 3], scratch);
        for (uint32_t j = 0; j < nt; ++j) {
            Tile t;
            t.start = row.off + uint64_t(j) * a.cfg.tile_len;
            t.v = v;
            t.len = min(a.cfg.tile_len, row.deg - j * a.cfg.tile_len);
            t.payload = __float_as_uint(d);
            t.pad = 0;
            tiles[t0 + j] = t;
        }
        if (nt) {
            ctr.scanned += row.deg;
            if (row.off & kHostTier) ctr.host += row.deg;
            d = 0.0f;
        }
        if (i < nq) a.delta[i] = d;
    }
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
}

// PageRank round, scatter phase (pr_residual.cuh:74-100); the closing
// workgroup advances the round.
template <int QPL>
__global__ __launch_bounds__(kBlock) void k_pr_scatter(PrArgs a, uint64_t *prog, uint64_t tag) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ RoundStage stage;
    Ctl *c = a.ctl;
    const uint32_t r = c->round;
    const uint32_t nq = c->qcount[r % 3];
    if (nq == 0) return post_idle(prog, tag);
    const uint32_t nt = c->tcount[r % 3];
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    Tile *tiles = (r & 1) ? a.t1 : a.t0;
    PrScatterOp op{a.rows, a.res, a.queued, a.delta, &c->pr_invalid, a.eps};
    // tiles of long rows were made while consuming; the scatter makes none
    const StagedOut<2048, 128> out{Sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]},
                                   TileSink{tiles, &c->scount}, &stage};
    out_init(out);
    Counters ctr;
    expand_range<QPL>(op, a.cfg, qin, nq, out, lds[wave_in_block()], ctr);
    expand_tiles<QPL>(op, a.cfg, tiles, nt, out);
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
    close_round(c, [&] {
        c->qcount[r % 3] = 0;
        c->tcount[r % 3] = 0;
        c->round = r + 1;
        c->rounds += 1;
        c->converged = ld_acquire(&c->qcount[(r + 1) % 3]) == 0 ? 1u : 0u;
        post_progress(prog, tag, c->converged == 0u);
    });
}

// Plain streaming copy (16 B per lane per step): the HBM reference point for
// rocprofv3 FETCH_SIZE / WRITE_SIZE calibration and achievable bandwidth.
__global__ void k_copy_probe(const uint4 *src, uint4 *dst, uint64_t n16) {
    for (uint64_t i = gthread(); i < n16; i += gthreads()) dst[i] = src[i];
}

// FP64 residual norm sum |r| (the only FP64 arithmetic of the PR path).
__global__ void k_pr_l1(const float *res, uint32_t n, StatShard *stats) {
    double s = 0;
    for (uint32_t v = gthread(); v < n; v += gthreads()) s += fabs(double(res[v]));
    for (int d = kWave / 2; d > 0; d >>= 1) s += __shfl_xor(s, d, kWave);
    if (lane_id() == 0 && s != 0) atomicAdd(&stat_shard(stats)->l1, s);
}

// Folds the statistics shards into the control block (adding to its
// totals) and clears them. One workgroup of 64 threads: lane j reads shard j.
__global__ void k_reduce_stats(StatShard *stats, Ctl *c) {
    static_assert(kStatShards == kWave, "one lane per shard");
    const uint32_t j = threadIdx.x;
    uint64_t *w = reinterpret_cast<uint64_t *>(stats + j);
    uint64_t v[kStatWords];
#pragma unroll
    for (uint32_t f = 0; f < kStatWords; ++f) {
        v[f] = wave_sum64(w[f]);
        w[f] = 0;
    }
    double l1 = stats[j].l1;
    stats[j].l1 = 0;
    for (int d = kWave / 2; d > 0; d >>= 1) l1 += __shfl_xor(l1, d, kWave);
    if (j == 0) {
        for (int i = 0; i < 4; ++i) {
            c->st_edges[i] += v[i];
            c->st_vexp[i] += v[4 + i];
            c->st_host[i] += v[8 + i];
            c->edges_scanned += v[i];
            c->vexp += v[4 + i];
            c->host_edges += v[8 + i];
        }
        c->deletes_applied += v[12];
        c->missing += v[13];
        c->invalid += v[14];
        c->adds_applied += v[15];
        c->relocated += v[16];
        c->garbage[0] += v[17];
        c->garbage[1] += v[18];
        c->residual_l1 += l1;
    }
}

// Pool compaction: copies every row of `tier` to its new offset.
__global__ void k_relocate_rows(const Row *old_rows, Row *new_rows, uint32_t n, uint32_t tier,
                                const uint32_t *src_pool, uint32_t *dst_pool) {
    const uint32_t lane = lane_id();
    for (uint32_t v = gwave(); v < n; v += gwaves()) {
        const Row o = old_rows[v];
        const bool in_tier = ((o.off & kHostTier) != 0) == (tier == 1);
        if (!in_tier || o.deg == 0) continue;
        const Row nr = new_rows[v];
        const uint32_t *s = src_pool + (o.off & ~kHostTier);
        uint32_t *d = dst_pool + (nr.off & ~kHostTier);
        for (uint32_t q = lane; q < o.deg; q += kWave) d[q] = s[q];
    }
}

// Moves rows between tiers (cache admission / eviction). Each job names a row
// whose descriptor has already been rewritten; `from` is the old location.
__global__ void k_move_rows(const CopyJob *jobs, const uint32_t *njobs, Pools pools) {
    const uint32_t cnt = ld_acquire(njobs);
    const uint32_t lane = lane_id();
    for (uint32_t j = gwave(); j < cnt; j += gwaves()) {
        const CopyJob jb = jobs[j];
        const uint32_t *s = pools.row(jb.from);
        uint32_t *d = pools.row(jb.to);
        for (uint32_t q = lane; q < jb.len; q += kWave) d[q] = s[q];
    }
}

// Activity window (comp_hotness_sssp / reset_hotness, driver.cuh:462-521):
// history keeps four 8-bit saturating window counts; score = their sum.
__global__ void k_age_hotness(uint32_t *hot, uint32_t *hist, uint32_t *score, uint32_t n) {
    for (uint32_t v = gthread(); v < n; v += gthreads()) {
        const uint32_t h = min(hot[v], 255u);
        const uint32_t nh = (hist[v] << 8) | h;
        hist[v] = nh;
        hot[v] = 0;
        score[v] = (nh & 255u) + ((nh >> 8) & 255u) + ((nh >> 16) & 255u) + (nh >> 24);
    }
}

}  // namespace dev
}  // namespace cig
