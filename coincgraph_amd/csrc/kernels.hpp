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
    CopyJob *jobs;
};

// GroupedUpdateBatch (effective_update_batch.h:35-87) on the device: count
// records per source and collect the touched sources.
__global__ void k_count(MutArgs a) {
    const uint32_t total = a.nd + a.na;
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
        const uint64_t badm = __ballot(bad);
        if (badm && lane_id() == 0)
            atomicAdd((unsigned long long *)&a.ctl->invalid, (unsigned long long)__popcll(badm));
        wave_append(first, u, a.touched, &a.ctl->touched);
    }
}

// Reserves each touched source's group and classifies its work.
__global__ void k_plan(MutArgs a) {
    const uint32_t touched = ld_acquire(&a.ctl->touched);
    for (uint32_t base = blockIdx.x * blockDim.x; base < touched; base += gthreads()) {
        const uint32_t t = base + threadIdx.x;
        bool small = false, big = false, add = false;
        uint32_t u = 0, len = 0, nd = 0, na = 0;
        if (t < touched) {
            u = a.touched[t];
            const uint64_t c = a.cnt64[u];
            nd = uint32_t(c);
            na = uint32_t(c >> 32);
            len = nd + na;
            // one wave scans a row of up to 4096 slots; longer rows (or more than
            // 64 deletes) get a whole workgroup
            small = nd != 0 && nd <= 64 && a.rows[u].deg <= 4096;
            big = nd != 0 && !small;
            add = na != 0;
        }
        // one group-cursor atomic per wave
        const uint32_t incl = wave_incl_scan(len);
        const uint32_t total_len = __shfl(incl, kWave - 1, kWave);
        unsigned long long gbase = 0;
        if (lane_id() == 0 && total_len)
            gbase = atomicAdd((unsigned long long *)&a.ctl->group_cursor, (unsigned long long)total_len);
        gbase = __shfl(gbase, 0, kWave);
        if (t < touched) a.gofs[u] = gbase + (incl - len);
        wave_append(small, u, a.small_list, &a.ctl->small_cnt);
        wave_append(big, u, a.big_list, &a.ctl->big_cnt);
        wave_append(add, u, a.add_list, &a.ctl->add_cnt);
    }
}

__global__ void k_scatter(MutArgs a) {
    const uint32_t total = a.nd + a.na;
    for (uint32_t i = gthread(); i < total; i += gthreads()) {
        const bool del = i < a.nd;
        const cig_edge e = del ? a.dels[i] : a.adds[i - a.nd];
        if (e.u >= a.n || e.v >= a.n) continue;
        const uint64_t pos = atomicAdd((unsigned long long *)&a.fill64[e.u],
                                       del ? 1ull : (1ull << 32));
        const uint32_t slot = del ? uint32_t(pos) : uint32_t(a.cnt64[e.u]) + uint32_t(pos >> 32);
        a.grouped[a.gofs[e.u] + slot] = e.v;
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
    if (lane == 0) {
        if (acc_applied) atomicAdd((unsigned long long *)&a.ctl->deletes_applied, (unsigned long long)acc_applied);
        if (acc_missing) atomicAdd((unsigned long long *)&a.ctl->missing, (unsigned long long)acc_missing);
    }
}

// Deletes for long rows or many deletes, one 1024-thread workgroup per row:
// the delete multiset goes to an LDS hash table (up to kPass deletes per
// pass); the row is scanned in parallel, each match claims one requested
// occurrence; holes below the new length are refilled from the live tail.
// The multiset of survivors equals the reference's; their order may differ.
constexpr int kDelBigThreads = 1024;
constexpr uint32_t kDelHT = 4096;
constexpr uint32_t kDelPass = 2048;

__device__ __forceinline__ uint32_t ht_hash(uint32_t x, uint32_t bits) {
    return (x * 0x9E3779B1u) >> (32 - bits);
}

__global__ __launch_bounds__(kDelBigThreads) void k_del_big(MutArgs a) {
    __shared__ uint32_t hkey[kDelHT];
    __shared__ int32_t hcnt[kDelHT];
    __shared__ uint32_t holes[kDelPass];
    __shared__ uint32_t fh[kDelPass];
    __shared__ uint32_t dead[kDelPass / 32];
    __shared__ uint32_t s_nh, s_fc, s_lc;
    const uint32_t cnt = ld_acquire(&a.ctl->big_cnt);
    const uint32_t tid = threadIdx.x;
    for (uint32_t j = blockIdx.x; j < cnt; j += gridDim.x) {
        const uint32_t u = a.big_list[j];
        const Row r = a.rows[u];
        const uint32_t nd = uint32_t(a.cnt64[u]);
        const uint32_t *dl = a.grouped + a.gofs[u];
        uint32_t *row = a.pools.row(r.off);
        uint32_t deg = r.deg, total = 0;
        for (uint32_t p0 = 0; p0 < nd; p0 += kDelPass) {
            const uint32_t pn = min(kDelPass, nd - p0);
            // table of 2*pn slots rounded up to a power of two (64..4096)
            uint32_t bits = 6;
            while ((1u << bits) < 2 * pn) ++bits;
            const uint32_t ht = 1u << bits;
            for (uint32_t s = tid; s < ht; s += kDelBigThreads) { hkey[s] = kNone; hcnt[s] = 0; }
            for (uint32_t s = tid; s < kDelPass / 32; s += kDelBigThreads) dead[s] = 0;
            if (tid == 0) { s_nh = 0; s_fc = 0; s_lc = 0; }
            __syncthreads();
            for (uint32_t q = tid; q < pn; q += kDelBigThreads) {
                const uint32_t d = dl[p0 + q];
                uint32_t h = ht_hash(d, bits);
                for (;;) {
                    const uint32_t prev = atomicCAS(&hkey[h], kNone, d);
                    if (prev == kNone || prev == d) { atomicAdd(&hcnt[h], 1); break; }
                    h = (h + 1) & (ht - 1);
                }
            }
            __syncthreads();
            for (uint32_t b = 0; b < deg; b += kDelBigThreads * 4) {
                uint32_t x[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t pos = b + uint32_t(k) * kDelBigThreads + tid;
                    x[k] = pos < deg ? row[pos] : kNone;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t pos = b + uint32_t(k) * kDelBigThreads + tid;
                    if (pos >= deg) continue;
                    uint32_t h = ht_hash(x[k], bits);
                    while (hkey[h] != kNone && hkey[h] != x[k]) h = (h + 1) & (ht - 1);
                    if (hkey[h] == x[k] && atomicSub(&hcnt[h], 1) > 0)
                        holes[atomicAdd(&s_nh, 1u)] = pos;
                }
            }
            __syncthreads();
            const uint32_t k = s_nh;
            const uint32_t nd2 = deg - k;
            for (uint32_t q = tid; q < k; q += kDelBigThreads) {
                const uint32_t p = holes[q];
                if (p < nd2) fh[atomicAdd(&s_fc, 1u)] = p;
                else atomicOr(&dead[(p - nd2) >> 5], 1u << ((p - nd2) & 31));
            }
            __syncthreads();
            for (uint32_t q = tid; q < k; q += kDelBigThreads) {
                if ((dead[q >> 5] >> (q & 31)) & 1u) continue;
                const uint32_t val = row[nd2 + q];
                row[fh[atomicAdd(&s_lc, 1u)]] = val;
            }
            __threadfence();
            __syncthreads();
            deg = nd2;
            total += k;
        }
        if (tid == 0) {
            a.rows[u].deg = deg;
            if (total) atomicAdd((unsigned long long *)&a.ctl->deletes_applied, (unsigned long long)total);
            if (nd - total) atomicAdd((unsigned long long *)&a.ctl->missing, (unsigned long long)(nd - total));
        }
        __syncthreads();
    }
}

// Appends the adds of each touched source (one wave per row); a full row
// moves to a block of twice the capacity at the tail of its tier
// (PatchOrInvalidateCachedAdjacency's grow path, csr_graph.cuh:1793-1821, and
// RewriteSourceToNewBlock, source_local_chunk_store.h:782-848). A row that
// cannot be placed is listed in `overflow` and left for the host to retry.
__global__ void k_add(MutArgs a, const uint32_t *list, const uint32_t *list_count) {
    const uint32_t cnt = ld_acquire(list_count);
    const uint32_t lane = lane_id();
    uint64_t acc_adds = 0, acc_reloc = 0, acc_garbage[2] = {0, 0};
    for (uint32_t j = gwave(); j < cnt; j += gwaves()) {
        const uint32_t u = list[j];
        const uint64_t c = a.cnt64[u];
        const uint32_t nd = uint32_t(c), na = uint32_t(c >> 32);
        const uint32_t *al = a.grouped + a.gofs[u] + nd;
        const Row r = a.rows[u];
        const uint64_t newdeg = uint64_t(r.deg) + na;
        if (newdeg <= r.cap) {
            uint32_t *row = a.pools.row(r.off);
            for (uint32_t q = lane; q < na; q += kWave) row[r.deg + q] = al[q];
            if (lane == 0) a.rows[u].deg = uint32_t(newdeg);
        } else {
            const uint32_t tier = (r.off & kHostTier) ? 1u : 0u;
            const uint64_t newcap = next_pow2_dev(newdeg);
            unsigned long long o = 0;
            if (lane == 0) o = atomicAdd((unsigned long long *)&a.ctl->pool_tail[tier], (unsigned long long)newcap);
            o = __shfl(o, 0, kWave);
            if (o + newcap > a.ctl->pool_cap[tier] || newcap > 0xFFFFFFFFull) {
                if (lane == 0) a.overflow[atomicAdd(&a.ctl->overflow_cnt, 1u)] = u;
                continue;
            }
            const uint64_t noff = uint64_t(o) | (tier ? kHostTier : 0ull);
            const uint32_t *src = a.pools.row(r.off);
            uint32_t *dst = a.pools.row(noff);
            if (r.deg <= 1024) {
                constexpr int U = 8;
                for (uint32_t b = 0; b < r.deg; b += kWave * U) {
                    uint32_t v[U];
#pragma unroll
                    for (int q = 0; q < U; ++q) {
                        const uint32_t p = b + uint32_t(q) * kWave + lane;
                        v[q] = p < r.deg ? src[p] : 0u;
                    }
#pragma unroll
                    for (int q = 0; q < U; ++q) {
                        const uint32_t p = b + uint32_t(q) * kWave + lane;
                        if (p < r.deg) dst[p] = v[q];
                    }
                }
            } else if (lane == 0) {
                CopyJob jb;
                jb.from = r.off;
                jb.to = noff;
                jb.len = r.deg;
                jb.pad = 0;
                a.jobs[atomicAdd(&a.ctl->copy_cnt, 1u)] = jb;
            }
            for (uint32_t q = lane; q < na; q += kWave) dst[r.deg + q] = al[q];
            if (lane == 0) {
                Row nr;
                nr.off = noff;
                nr.deg = uint32_t(newdeg);
                nr.cap = uint32_t(newcap);
                a.rows[u] = nr;
            }
            acc_garbage[tier] += r.cap;
            acc_reloc += 1;
        }
        acc_adds += na;
    }
    if (lane == 0) {
        if (acc_adds) atomicAdd((unsigned long long *)&a.ctl->adds_applied, (unsigned long long)acc_adds);
        if (acc_reloc) atomicAdd((unsigned long long *)&a.ctl->relocated, (unsigned long long)acc_reloc);
        for (int t = 0; t < 2; ++t)
            if (acc_garbage[t]) atomicAdd((unsigned long long *)&a.ctl->garbage[t], (unsigned long long)acc_garbage[t]);
    }
}

// Copies long relocated rows: every workgroup takes 1024-slot tiles of each
// job, so one long row is spread over many workgroups.
__global__ void k_copy_jobs(MutArgs a) {
    const uint32_t cnt = ld_acquire(&a.ctl->copy_cnt);
    constexpr uint32_t kTile = 1024;
    for (uint32_t j = 0; j < cnt; ++j) {
        const CopyJob jb = a.jobs[j];
        const uint32_t *src = a.pools.row(jb.from);
        uint32_t *dst = a.pools.row(jb.to);
        for (uint32_t t = blockIdx.x * kTile; t < jb.len; t += gridDim.x * kTile) {
            uint32_t v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t p = t + uint32_t(q) * blockDim.x + threadIdx.x;
                v[q] = p < jb.len ? src[p] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t p = t + uint32_t(q) * blockDim.x + threadIdx.x;
                if (p < jb.len) dst[p] = v[q];
            }
        }
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
        wave_append(hit, v, a.alist, &a.ctl->acount);
    }
}

// Invalidation waves (RunSyncPushDDB_del rounds, framework.cuh:2815-2853) as
// one persistent launch: the affected list grows in place; round k expands
// the items appended by round k-1.
template <int EPL>
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
        Sink sink{a.alist, &c->acount};
        TileSink ts{(k & 1) ? a.t0 : a.t1, &c->icount[(k + 1) % 3]};
        expand_range<EPL>(op, a.cfg, a.alist + b, e - b, sink, ts, L, ctr);
        expand_tiles<EPL>(op, a.cfg, (k & 1) ? a.t1 : a.t0, nt, sink, ctr);
        b = e;
        cg::this_grid().sync();
    }
    flush_counters(ctr, c, a.cfg.stage);
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
template <int EPL>
__global__ __launch_bounds__(kBlock) void k_inval_round(InvalArgs a) {
    __shared__ WaveLds lds[kWavesPerBlock];
    Ctl *c = a.ctl;
    const uint32_t b = c->ibegin, e = c->iend, k = c->iround;
    const uint32_t nt = c->icount[k % 3];
    if (b == e && nt == 0) return;
    InvalidateOp op{a.rows, a.state, a.aff};
    Counters ctr;
    Sink sink{a.alist, &c->acount};
    TileSink ts{(k & 1) ? a.t0 : a.t1, &c->icount[(k + 1) % 3]};
    expand_range<EPL>(op, a.cfg, a.alist + b, e - b, sink, ts, lds[wave_in_block()], ctr);
    expand_tiles<EPL>(op, a.cfg, (k & 1) ? a.t1 : a.t0, nt, sink, ctr);
    flush_counters(ctr, c, a.cfg.stage);
    close_round(c, [&] {
        c->icount[k % 3] = 0;
        c->ibegin = e;
        c->iend = ld_acquire(&c->acount);
        c->iround = k + 1;
    });
}

// reset_affected_values (driver.cuh:418-433): affected vertices restart from
// the initial value (CC: own label), and all of them form the next frontier.
__global__ void k_reset_affected(InvalArgs a) {
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
        wave_append(push, v, q, &a.ctl->qcount[r % 3]);
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
        wave_append(push, v, q, &a.ctl->qcount[r % 3]);
    }
}

// -------------------------------------------------------------- repair scan
// Every workgroup first folds the affected list into an LDS bitmap filter of
// 2^18 bits (a multiplicative hash of the id); an edge whose destination
// misses the filter is not affected, so almost every edge of the scan is
// settled in LDS and only filter hits read the global affected bitmap. The
// filter is skipped when the affected list is too long to stay selective.
constexpr uint32_t kFilterBits = 18;
constexpr uint32_t kFilterWords = (1u << kFilterBits) / 32;
constexpr uint32_t kFilterMaxItems = 1u << 16;

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

__device__ bool build_filter(uint32_t *lds, const uint32_t *alist, uint32_t count) {
    if (count > kFilterMaxItems) return false;
    for (uint32_t w = threadIdx.x; w < kFilterWords; w += blockDim.x) lds[w] = 0u;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
        const uint32_t h = filter_hash(alist[i]);
        atomicOr(&lds[h >> 5], 1u << (h & 31));
    }
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

template <int W, bool kStrict, int EPL>
__global__ __launch_bounds__(kBlock) void k_repair_scan(InvalArgs a) {
    __shared__ uint32_t filt[kFilterWords];
    __shared__ WaveLds lds[kWavesPerBlock];
    Ctl *c = a.ctl;
    const uint32_t na = ld_acquire(&c->acount);
    if (na == 0) return;
    const bool on = build_filter(filt, a.alist, na);
    RepairScanOp<W, kStrict> op{a.rows, a.state, AffFilter{on ? filt : nullptr, a.aff}, a.seed};
    Counters ctr;
    Sink none{nullptr, &c->scount};
    TileSink ts{a.t0, &c->scount};
    expand_range<EPL>(op, a.cfg, nullptr, a.n, none, ts, lds[wave_in_block()], ctr);
    flush_counters(ctr, c, a.cfg.stage);
}

template <int W, bool kStrict, int EPL>
__global__ __launch_bounds__(kBlock) void k_repair_tiles(InvalArgs a) {
    __shared__ uint32_t filt[kFilterWords];
    Ctl *c = a.ctl;
    const uint32_t na = ld_acquire(&c->acount);
    const uint32_t nt = ld_acquire(&c->scount);
    if (na == 0 || nt == 0) return;
    const bool on = build_filter(filt, a.alist, na);
    RepairScanOp<W, kStrict> op{a.rows, a.state, AffFilter{on ? filt : nullptr, a.aff}, a.seed};
    Counters ctr;
    Sink none{nullptr, &c->scount};
    expand_tiles<EPL>(op, a.cfg, a.t0, nt, none, ctr);
}

// ------------------------------------------------- one-shot expansion passes
template <int EPL, class Op>
__global__ __launch_bounds__(kBlock) void k_expand_once(Op op, ExpandCfg cfg, const uint32_t *items,
                                                         const uint32_t *count_ptr, uint32_t count,
                                                         Sink sink, TileSink ts, Ctl *ctl,
                                                         const uint32_t *gate) {
    __shared__ WaveLds lds[kWavesPerBlock];
    if (gate && ld_acquire(gate) == 0) return;
    Counters ctr;
    const uint32_t n = count_ptr ? ld_acquire(count_ptr) : count;
    expand_range<EPL>(op, cfg, items, n, sink, ts, lds[wave_in_block()], ctr);
    flush_counters(ctr, ctl, cfg.stage);
}

template <int EPL, class Op>
__global__ __launch_bounds__(kBlock) void k_expand_tiles(Op op, ExpandCfg cfg, const Tile *tiles,
                                                          const uint32_t *ntiles, Sink sink, Ctl *ctl) {
    Counters ctr;
    expand_tiles<EPL>(op, cfg, tiles, ld_acquire(ntiles), sink, ctr);
    flush_counters(ctr, ctl, cfg.stage);
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
template <int W, bool kStrict, int EPL>
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
        Sink sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]};
        TileSink ts{(r & 1) ? a.t0 : a.t1, &c->tcount[(r + 1) % 3]};
        expand_range<EPL>(op, a.cfg, qin, nq, sink, ts, L, ctr);
        expand_tiles<EPL>(op, a.cfg, (r & 1) ? a.t1 : a.t0, nt, sink, ctr);
        cg::this_grid().sync();
    }
    flush_counters(ctr, c, a.cfg.stage);
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
// enqueues rounds back to back; a round whose frontier and deferred tiles
// are empty returns at once.
template <int W, bool kStrict, int EPL>
__global__ __launch_bounds__(kBlock) void k_relax_round(LoopArgs a) {
    __shared__ WaveLds lds[kWavesPerBlock];
    Ctl *c = a.ctl;
    const uint32_t r = c->round;
    const uint32_t nq = c->qcount[r % 3], nt = c->tcount[r % 3];
    if (nq == 0 && nt == 0) return;
    Counters ctr;
    RelaxOp<W, kStrict> op{a.rows, a.state, a.stamp, a.hot, a.seed, r + 1};
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    Sink sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]};
    TileSink ts{(r & 1) ? a.t0 : a.t1, &c->tcount[(r + 1) % 3]};
    expand_range<EPL>(op, a.cfg, qin, nq, sink, ts, lds[wave_in_block()], ctr);
    expand_tiles<EPL>(op, a.cfg, (r & 1) ? a.t1 : a.t0, nt, sink, ctr);
    flush_counters(ctr, c, a.cfg.stage);
    close_round(c, [&] {
        c->qcount[r % 3] = 0;
        c->tcount[r % 3] = 0;
        c->round = r + 1;
        c->rounds += 1;
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
template <int EPL>
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
        Sink sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]};
        TileSink ts{tiles, &c->scount};  // unused: long rows were tiled while consuming
        Counters sctr;
        expand_range<EPL>(op, a.cfg, qin, nq, sink, ts, L, sctr);
        ctr.scanned += sctr.scanned;
        ctr.host += sctr.host;
        expand_tiles<EPL>(op, a.cfg, tiles, ld_acquire(&c->tcount[r % 3]), sink, ctr);
        cg::this_grid().sync();
    }
    flush_counters(ctr, c, a.cfg.stage);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->round = r;
        c->rounds += it;
        c->converged = ld_acquire(&c->qcount[r % 3]) == 0 ? 1u : 0u;
    }
}

// PageRank round, consume phase (pr_residual.cuh:54-71) as an ordinary
// launch; long rows are cut into tiles for this round's scatter.
__global__ __launch_bounds__(kBlock) void k_pr_consume(PrArgs a) {
    Ctl *c = a.ctl;
    const uint32_t r = c->round;
    const uint32_t nq = c->qcount[r % 3];
    if (nq == 0) return;
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    Tile *tiles = (r & 1) ? a.t1 : a.t0;
    Counters ctr;
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
    flush_counters(ctr, c, a.cfg.stage);
}

// PageRank round, scatter phase (pr_residual.cuh:74-100); the closing
// workgroup advances the round.
template <int EPL>
__global__ __launch_bounds__(kBlock) void k_pr_scatter(PrArgs a) {
    __shared__ WaveLds lds[kWavesPerBlock];
    Ctl *c = a.ctl;
    const uint32_t r = c->round;
    const uint32_t nq = c->qcount[r % 3];
    if (nq == 0) return;
    const uint32_t nt = c->tcount[r % 3];
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    Tile *tiles = (r & 1) ? a.t1 : a.t0;
    PrScatterOp op{a.rows, a.res, a.queued, a.delta, &c->pr_invalid, a.eps};
    Sink sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]};
    TileSink ts{tiles, &c->scount};  // unused: long rows were tiled while consuming
    Counters ctr;
    expand_range<EPL>(op, a.cfg, qin, nq, sink, ts, lds[wave_in_block()], ctr);
    expand_tiles<EPL>(op, a.cfg, tiles, nt, sink, ctr);
    flush_counters(ctr, c, a.cfg.stage);
    close_round(c, [&] {
        c->qcount[r % 3] = 0;
        c->tcount[r % 3] = 0;
        c->round = r + 1;
        c->rounds += 1;
        c->converged = ld_acquire(&c->qcount[(r + 1) % 3]) == 0 ? 1u : 0u;
    });
}

// Plain streaming copy (16 B per lane per step): the HBM reference point for
// rocprofv3 FETCH_SIZE / WRITE_SIZE calibration and achievable bandwidth.
__global__ void k_copy_probe(const uint4 *src, uint4 *dst, uint64_t n16) {
    for (uint64_t i = gthread(); i < n16; i += gthreads()) dst[i] = src[i];
}

// FP64 residual norm sum |r| (the only FP64 arithmetic of the PR path).
__global__ void k_pr_l1(const float *res, uint32_t n, Ctl *c) {
    double s = 0;
    for (uint32_t v = gthread(); v < n; v += gthreads()) s += fabs(double(res[v]));
    for (int d = kWave / 2; d > 0; d >>= 1) s += __shfl_xor(s, d, kWave);
    if (lane_id() == 0 && s != 0) atomicAdd(&c->residual_l1, s);
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
