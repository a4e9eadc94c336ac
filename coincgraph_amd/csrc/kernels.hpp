// Kernels of the incremental graph engine. Included once, by engine.hip.
#pragma once
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "../../include/cig.h"
#include "device.hpp"

namespace cig {
namespace dev {

namespace cg = cooperative_groups;

// LDS staging of the round kernels: 2048 frontier items and 128 tiles per
// workgroup (19 KB with the wave slots of a 256-thread workgroup).
using RoundStage = StageLds<2048, 128>;

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
// flags: kResetKeepRounds keeps the round state; kResetKeepMut keeps the
// mutation counters (the second stream resets those itself, k_reset_mut).
constexpr uint32_t kResetKeepRounds = 1u, kResetKeepMut = 2u;
__global__ void k_reset_ctl(Ctl *c, uint32_t flags) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const bool keep_rounds = flags & kResetKeepRounds;
    c->acount = 0;
    c->icount[0] = c->icount[1] = c->icount[2] = 0;
    c->ifront[0] = c->ifront[1] = c->ifront[2] = 0;
    c->iround = 0;
    c->scount = 0;
    if (!(flags & kResetKeepMut)) {
        c->touched = 0;
        c->small_cnt = c->big_cnt = c->add_cnt = c->copy_cnt = c->overflow_cnt = 0;
        c->mid_cnt = c->mid_host_cnt = c->mid16_cnt = 0;
        c->dtile_cnt = c->ht_cursor = 0;
        c->group_cursor = 0;
    }
    c->pr_invalid = 0;
    c->pull_rounds = 0;
    c->edges_scanned = c->host_edges = c->vexp = 0;
    c->deletes_applied = c->missing = c->invalid = c->adds_applied = c->relocated = 0;
    c->touched_sources = 0;
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
    c->mid_cnt = c->mid_host_cnt = c->mid16_cnt = 0;
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
// A batch's records in device memory. Launches replayed from a captured
// HIP graph read the records through it (the graph's arguments stay the
// same for every batch; a one-thread kernel before each replay fills it).
struct BatchRef {
    const cig_edge *dels;
    const cig_edge *adds;
    uint32_t nd, na;
};

__global__ void k_set_batch(BatchRef *ref, const cig_edge *dels, uint32_t nd, const cig_edge *adds, uint32_t na) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *ref = BatchRef{dels, adds, nd, na};
}

// Update records from the caller's vertex ids to a renumbered engine's;
// ids >= n stay as they are, so invalid records are still rejected.
__global__ void k_map_records(cig_edge *r, uint64_t cnt, const uint32_t *perm, uint32_t n) {
    for (uint64_t i = gthread(); i < cnt; i += gthreads()) {
        cig_edge e = r[i];
        if (e.u < n) e.u = perm[e.u];
        if (e.v < n) e.v = perm[e.v];
        r[i] = e;
    }
}

// Freed blocks of the pinned host tier, one stack per size class (class =
// log2 of the block's slots). A relocated host-tier row's old block is pushed
// once its slots have been copied (k_fl_release, after k_add_copy; tier
// moves: k_move_release after k_move_rows), and a later relocation or
// eviction of the same class pops it instead of growing the tail. Pushes and
// pops never share a launch. A full stack drops the block, which stays
// garbage until the tier is packed; packing the tier empties every stack.
struct FreeList {
    unsigned long long *off;  // [kFlClasses * cap] block offsets (tier bit clear); null: no list
    int *cnt;                 // [kFlClasses]
    uint32_t cap;
};
constexpr uint32_t kFlClasses = 32;

__device__ __forceinline__ uint32_t fl_class(uint64_t cap) { return 63u - uint32_t(__clzll(cap)); }

__device__ __forceinline__ bool fl_pop(const FreeList &f, uint32_t cls, uint64_t &off) {
    if (!f.off) return false;
    const int old = atomicSub(&f.cnt[cls], 1);
    if (old > 0) {
        off = f.off[uint64_t(cls) * f.cap + uint32_t(old - 1)];
        return true;
    }
    atomicAdd(&f.cnt[cls], 1);  // empty: undo (no push runs in the same launch)
    return false;
}

__device__ __forceinline__ void fl_push(const FreeList &f, uint32_t cls, uint64_t off) {
    if (!f.off) return;
    const int at = atomicAdd(&f.cnt[cls], 1);
    if (at < int(f.cap)) f.off[uint64_t(cls) * f.cap + uint32_t(at)] = off;
    else atomicSub(&f.cnt[cls], 1);  // full: the block stays garbage
}

struct MutArgs {
    Ctl *ctl;
    Row *rows;
    Pools pools;
    FreeList fl;       // host-tier free blocks (see FreeList)
    uint32_t n;
    uint32_t rev;      // 1: rows [n, 2n) hold the incoming rows and get every update mirrored
    const cig_edge *dels;
    uint32_t nd;
    const cig_edge *adds;
    uint32_t na;
    uint64_t *cnt64;   // per vertex: deletes (low) | adds (high) of this call
    uint32_t *rslot;   // per record: its rank among its row's deletes (or adds), from k_count
    uint64_t *gofs;    // per vertex offset of its group in `grouped`
    uint32_t *touched;
    uint32_t *grouped;
    uint32_t *small_list;
    uint32_t *mid_list;   // medium rows with deletes (k_del_rows): HBM-tier rows from the
                          // front, host-tier rows from the back (n << rev entries)
    uint32_t *mid16_list; // HBM-tier medium rows with 9-16 deletes (n << rev entries)
    uint32_t *big_list;
    uint32_t *add_list;
    uint32_t *overflow;
    // long-row deletes
    uint32_t *htoff;   // per vertex: offset of its delete table (kNone: none)
    uint32_t *hkeys;   // delete tables: destination + 1 (0 = empty)
    uint2 *hcnts;      // delete tables: (requested, claimed)
    uint32_t *hcnt;    // per vertex: holes found (zero between calls)
    uint8_t *dclass;   // per row: delete class of this call (kDelClass*)
    const uint8_t *skip;  // grouped batches: 1 = record cancelled (adds [0,na), deletes [na,na+nd)); may be null
    struct DelTile *dtiles;  // long-row delete scan tiles (DelTile below)
    struct AddPiece *pieces;  // copy pieces of the insert phase (AddPiece below)
    StatShard *stats;
    uint32_t *hot;     // per row activity of the current window (null: not counted)
    const BatchRef *ref;  // non-null: the records are read from here (graph replays)
};

// The call's records: from the batch reference when there is one.
template <class A>
__device__ __forceinline__ uint32_t rec_nd(const A &a) { return a.ref ? a.ref->nd : a.nd; }
template <class A>
__device__ __forceinline__ uint32_t rec_na(const A &a) { return a.ref ? a.ref->na : a.na; }
template <class A>
__device__ __forceinline__ const cig_edge *rec_dels(const A &a) { return a.ref ? a.ref->dels : a.dels; }
template <class A>
__device__ __forceinline__ const cig_edge *rec_adds(const A &a) { return a.ref ? a.ref->adds : a.adds; }

// Deletes find their slot by row length (BuildDeletionPlan,
// source_local_chunk_store.h:670-723, restated for the device):
//  - tiny rows (<= 64 slots): one lane per delete record reads the whole row
//    and claims a matching slot with a CAS (slot := kNone);
//  - medium rows (<= 4096 slots, <= kDelWaveMax deletes; HBM rows with up to
//    kDelWaveMax16 in a second kernel): a 16-lane group per row against all
//    of the row's deletes at once, 256 slots a step;
//  - other rows: tiles of kDelTile slots over the whole grid against the
//    row's delete multiset in an open-addressing table of 2^bits >= 2*nd
//    slots.
// Every claimed slot is recorded in the row's hole list (the row's group,
// whose delete records are not read again). Rows with at most 16 deletes
// (and no table) are then compacted by one thread, the others by a
// workgroup: live slots of the last k positions move into the holes below
// deg-k. Equal destinations are interchangeable, so the survivor multiset is
// the reference's; their order may differ (CompactSourceInPlace,
// source_local_chunk_store.h:733-780, keeps it).
constexpr uint32_t kDelTile = 4096;
constexpr uint32_t kDelLdsKeys = 2048;

// Everything a scan tile needs, written by k_plan so that the tile's row
// loads depend on one record read only.
struct DelTile {
    uint64_t off;   // pool offset of the tile's first slot (tier bit kept)
    uint32_t row, start, len, to, bits, pad;
};
static_assert(sizeof(DelTile) == 32, "DelTile layout");
constexpr uint32_t kDelTinyDeg = 64, kDelMediumDeg = 4096, kDelThreadFix = 16;
enum : uint8_t { kDelClassNone = 0, kDelClassTiny = 1, kDelClassMedium = 2, kDelClassLong = 3, kDelClassMedium16 = 4 };

// A medium row with many deletes is cheaper as one tile against its table
// than as one wave per delete (each of which may read the whole row).
constexpr uint32_t kDelWaveMax = 8;
// In large batches, HBM-tier medium rows with more deletes, up to this many,
// take a second medium-row pass whose groups hold 16 requests (one tile
// against the row's table costs far more for a row of a few thousand
// slots); the 8-request kernel keeps its registers. Batches of fewer delete
// records (mirrors included) than kMid16MinRecords have few such rows and
// keep the plain form, which saves them the extra classification and
// workgroups.
constexpr uint32_t kDelWaveMax16 = 16;
constexpr uint64_t kMid16MinRecords = 1u << 20;

template <bool M16>
__device__ __forceinline__ uint8_t del_class(uint32_t nd, uint32_t deg, bool host) {
    if (nd == 0) return kDelClassNone;
    if (deg <= kDelTinyDeg) return kDelClassTiny;
    if (deg > kDelMediumDeg) return kDelClassLong;
    if (nd <= kDelWaveMax) return kDelClassMedium;
    return M16 && nd <= kDelWaveMax16 && !host ? kDelClassMedium16 : kDelClassLong;
}

__device__ __forceinline__ uint32_t del_table_bits(uint32_t nd) {
    uint32_t bits = 4;
    while ((1ull << bits) < 2ull * nd) ++bits;
    return bits;
}

__device__ __forceinline__ uint32_t ht_hash(uint32_t x, uint32_t bits) {
    return (x * 0x9E3779B1u) >> (32 - bits);
}

// Record i of a mutation call: deletes then adds; with the incoming index on,
// records [nd+na, 2(nd+na)) are the same updates mirrored onto row n+v with
// destination u.
struct MutRecord {
    uint32_t row, dst;
    bool del, valid, bad, mirror;
};

__device__ __forceinline__ MutRecord mut_record(const MutArgs &a, uint32_t i) {
    const uint32_t nd = rec_nd(a), na = rec_na(a);
    const uint32_t total = nd + na;
    MutRecord r;
    r.mirror = i >= total;
    const uint32_t j = r.mirror ? i - total : i;
    r.del = j < nd;
    const cig_edge e = r.del ? rec_dels(a)[j] : rec_adds(a)[j - nd];
    r.bad = !(e.u < a.n && e.v < a.n);
    r.valid = !r.bad && !(a.skip && a.skip[r.del ? na + j : j - nd]);
    r.row = r.mirror ? a.n + e.v : e.u;
    r.dst = r.mirror ? e.u : e.v;
    return r;
}

__device__ __forceinline__ uint32_t mut_records(const MutArgs &a) { return (rec_nd(a) + rec_na(a)) << a.rev; }

// Delete record i of [0, nd << rev): the deletes, then their mirrors.
__device__ __forceinline__ MutRecord del_record(const MutArgs &a, uint32_t i) {
    const uint32_t nd = rec_nd(a);
    return mut_record(a, i < nd ? i : i - nd + nd + rec_na(a));
}

// GroupedUpdateBatch (effective_update_batch.h:35-87) on the device: count
// records per row and collect the touched rows. The count's atomic also
// ranks the record among its row's deletes (or adds), which k_scatter uses
// as its slot. A thread takes four records a step, so the touched-list
// reservation is one atomic per 1024 records.
__global__ void k_count(MutArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    constexpr uint32_t R = 4;
    const uint32_t total = mut_records(a);
    uint64_t bad_total = 0, fwd_touched = 0;
    for (uint32_t base = blockIdx.x * blockDim.x * R; base < total; base += gthreads() * R) {
        uint32_t u[R];
        bool first[R];
        uint32_t nf = 0;
#pragma unroll
        for (uint32_t k = 0; k < R; ++k) {
            const uint32_t i = base + k * blockDim.x + threadIdx.x;
            first[k] = false;
            u[k] = 0;
            bool bad = false;
            if (i < total) {
                const MutRecord r = mut_record(a, i);
                if (!r.valid) {
                    bad = r.bad && !r.mirror;
                } else {
                    u[k] = r.row;
                    const uint64_t old = atomicAdd((unsigned long long *)&a.cnt64[r.row],
                                                   r.del ? 1ull : (1ull << 32));
                    a.rslot[i] = r.del ? uint32_t(old) : uint32_t(old >> 32);
                    first[k] = old == 0;
                }
            }
            bad_total += uint32_t(__popcll(__ballot(bad)));
            fwd_touched += uint32_t(__popcll(__ballot(first[k] && u[k] < a.n)));
            nf += first[k] ? 1u : 0u;
        }
        uint32_t at = block_reserve<uint32_t>(nf, &a.ctl->touched, scratch);
#pragma unroll
        for (uint32_t k = 0; k < R; ++k)
            if (first[k]) a.touched[at++] = u[k];
    }
    StatShard *sh = stat_shard(a.stats);
    stat_add(&sh->invalid, bad_total);
    stat_add(&sh->touched, fwd_touched);
}

// Reserves each touched row's group and classifies its work; long rows
// with deletes also reserve their table and scan tiles. All cursors (eight,
// nine with M16: the 16-request medium rows of a large batch) are reserved
// together, one round trip per workgroup step.
template <bool M16>
__global__ void k_plan(MutArgs a) {
    constexpr int NC = M16 ? 9 : 8;
    __shared__ uint32_t scratch[NC * (kMaxWavesPerBlock + 1)];
    const uint32_t touched = ld_acquire(&a.ctl->touched);
    uint32_t *ctr[NC];
    ctr[0] = &a.ctl->group_cursor;
    ctr[1] = &a.ctl->ht_cursor;
    ctr[2] = &a.ctl->dtile_cnt;
    ctr[3] = &a.ctl->small_cnt;
    ctr[4] = &a.ctl->big_cnt;
    ctr[5] = &a.ctl->add_cnt;
    ctr[6] = &a.ctl->mid_cnt;
    ctr[7] = &a.ctl->mid_host_cnt;
    if constexpr (M16) ctr[8] = &a.ctl->mid16_cnt;
    for (uint32_t base = blockIdx.x * blockDim.x; base < touched; base += gthreads()) {
        const uint32_t t = base + threadIdx.x;
        uint32_t u = 0, nd = 0, na = 0, deg = 0;
        bool host = false;
        uint8_t cls = kDelClassNone;
        if (t < touched) {
            u = a.touched[t];
            const uint64_t c = a.cnt64[u];
            nd = uint32_t(c);
            na = uint32_t(c >> 32);
            const Row row = a.rows[u];
            deg = row.deg;
            host = (row.off & kHostTier) != 0;
            cls = del_class<M16>(nd, deg, host);
            a.dclass[u] = cls;
        }
        const bool table = cls == kDelClassLong;
        const bool thread_fix = nd != 0 && nd <= kDelThreadFix && !table;
        uint32_t x[NC];
        x[0] = nd + na;
        x[1] = table ? (1u << del_table_bits(nd)) : 0u;
        x[2] = table ? (deg + kDelTile - 1) / kDelTile : 0u;
        x[3] = thread_fix ? 1u : 0u;
        x[4] = (nd != 0 && !thread_fix) ? 1u : 0u;
        x[5] = na != 0 ? 1u : 0u;
        x[6] = cls == kDelClassMedium && !host ? 1u : 0u;
        x[7] = cls == kDelClassMedium && host ? 1u : 0u;
        if constexpr (M16) x[8] = cls == kDelClassMedium16 ? 1u : 0u;
        uint32_t at[NC];
        block_reserve_n<NC>(x, ctr, at, scratch);
        if (t < touched) {
            a.gofs[u] = at[0];
            a.htoff[u] = table ? at[1] : kNone;
            if (x[2]) {
                const uint64_t off = a.rows[u].off;
                const uint32_t bits = del_table_bits(nd);
                for (uint32_t j = 0; j < x[2]; ++j) {
                    DelTile dt;
                    dt.off = off + uint64_t(j) * kDelTile;
                    dt.row = u;
                    dt.start = j * kDelTile;
                    dt.len = min(kDelTile, deg - j * kDelTile);
                    dt.to = at[1];
                    dt.bits = bits;
                    dt.pad = 0;
                    a.dtiles[at[2] + j] = dt;
                }
            }
            if (x[3]) a.small_list[at[3]] = u;
            if (x[4]) a.big_list[at[4]] = u;
            if (x[5]) a.add_list[at[5]] = u;
            if (x[6]) a.mid_list[at[6]] = u;
            if (x[7]) a.mid_list[(a.n << a.rev) - 1u - at[7]] = u;
            if constexpr (M16)
                if (x[8]) a.mid16_list[at[8]] = u;
        }
    }
}

// Groups the records by source; deletes of long rows also enter their
// row's table (destination -> requested count).
__global__ void k_scatter(MutArgs a) {
    const uint32_t total = mut_records(a);
    for (uint32_t i = gthread(); i < total; i += gthreads()) {
        const MutRecord r = mut_record(a, i);
        if (!r.valid) continue;
        const uint32_t nd = uint32_t(a.cnt64[r.row]);
        const uint32_t slot = r.del ? a.rslot[i] : nd + a.rslot[i];
        a.grouped[a.gofs[r.row] + slot] = r.dst;
        const uint32_t to = r.del ? a.htoff[r.row] : kNone;
        if (to != kNone) {
            // at most nd distinct keys in 2^bits >= 2*nd slots: the probe ends
            const uint32_t bits = del_table_bits(nd), mask = (1u << bits) - 1u;
            uint32_t h = ht_hash(r.dst, bits);
            for (;;) {
                const uint32_t prev = atomicCAS(&a.hkeys[to + h], 0u, r.dst + 1u);
                if (prev == 0u || prev == r.dst + 1u) break;
                h = (h + 1u) & mask;
            }
            atomicAdd(&a.hcnts[to + h].x, 1u);
        }
    }
}

// Records a claimed slot in its row's hole list.
__device__ __forceinline__ void del_hole(const MutArgs &a, uint32_t row, uint32_t slot) {
    a.grouped[a.gofs[row] + atomicAdd(&a.hcnt[row], 1u)] = slot;
}

// Tiny rows: one lane per delete record.
__device__ void del_lane_body(const MutArgs &a, uint32_t bid, uint32_t nb) {
    const uint32_t total = rec_nd(a) << a.rev;
    for (uint32_t i = bid * blockDim.x + threadIdx.x; i < total; i += nb * blockDim.x) {
        const MutRecord r = del_record(a, i);
        if (!r.valid || a.dclass[r.row] != kDelClassTiny) continue;
        const Row row = a.rows[r.row];
        if (row.deg == 0) continue;  // nothing to delete: a missing record
        uint32_t *p = a.pools.row(row.off);
        const uint32_t nq = (row.deg + 3u) >> 2;  // 1..16
        uint32_t d[kDelTinyDeg / 4][4], live[kDelTinyDeg / 4];
#pragma unroll
        for (uint32_t q = 0; q < kDelTinyDeg / 4; ++q) {
            live[q] = load_quad(p, q < nq ? q : nq - 1, row.deg, d[q]);
            if (q >= nq) live[q] = 0;
        }
        // a row with one delete record has one claimant: a plain store
        // (for host-tier rows a CAS is an atomic over the host link)
        const bool sole = uint32_t(a.cnt64[r.row]) == 1u;
        bool done = false;
#pragma unroll
        for (uint32_t q = 0; q < kDelTinyDeg / 4; ++q)
#pragma unroll
            for (uint32_t s = 0; s < 4; ++s)
                if (!done && s < live[q] && d[q][s] == r.dst &&
                    (sole ? (p[4 * q + s] = kNone, true) : atomicCAS(&p[4 * q + s], r.dst, kNone) == r.dst)) {
                    del_hole(a, r.row, 4 * q + s);
                    done = true;
                }
    }
}

__global__ void k_del_lane(MutArgs a) { del_lane_body(a, blockIdx.x, gridDim.x); }

// Medium rows (64 < deg <= 4096 slots, at most K deletes: kDelWaveMax, or
// kDelWaveMax16 for the HBM rows of the second kernel): one
// 16-lane group per row (four rows per wave) reads the row once, 256 slots
// a step, against all of the row's requested destinations, which the
// group's lanes hold in registers (a destination requested twice is two
// requests). The group owns the row, so a slot equal to a destination with
// an open request is claimed with a plain store and its index goes to the
// row's hole list; the group stops as soon as every request is claimed. (A
// group per delete record would read the row once per request.) Q quads
// per lane a step: 4 for HBM rows; host-tier rows take 2 (HOST), since
// what a step reads past the last claim costs host-link bandwidth, which
// bounds them.
template <uint32_t Q, bool HOST, uint32_t K = kDelWaveMax>
__device__ void del_rows_body(const MutArgs &a, uint32_t bid, uint32_t nb) {
    constexpr uint32_t G = 16, GPW = kWave / G;
    static_assert(K <= G, "a group holds its row's requests");
    static_assert(!(HOST && K > kDelWaveMax), "the 16-request rows are HBM rows");
    constexpr bool M16 = K > kDelWaveMax;
    static_assert(4 * Q <= 32, "candidate bits fit one word");
    const uint32_t cnt = ld_acquire(M16 ? &a.ctl->mid16_cnt : HOST ? &a.ctl->mid_host_cnt : &a.ctl->mid_cnt);
    const uint32_t lane = lane_id(), g = lane / G, gl = lane % G;
    const uint64_t gmask = ((1ull << G) - 1ull) << (g * G);
    const uint32_t wpb = blockDim.x / kWave;
    for (uint32_t base = (bid * wpb + wave_in_block()) * GPW; base < cnt; base += nb * wpb * GPW) {
        const uint32_t i = base + g;
        uint32_t u = 0, nd = 0, nq = 0, deg = 0;
        uint32_t *grp = nullptr, *p = nullptr;
        if (i < cnt) {
            u = M16 ? a.mid16_list[i] : a.mid_list[HOST ? (a.n << a.rev) - 1u - i : i];
            nd = min(uint32_t(a.cnt64[u]), K);
            const Row row = a.rows[u];
            deg = row.deg;
            nq = (deg + 3u) >> 2;
            grp = a.grouped + a.gofs[u];  // the requests, then the hole list
            p = a.pools.row(row.off);
        }
        const uint32_t mine = gl < nd ? grp[gl] : kNone;
        uint32_t key[K];
#pragma unroll
        for (uint32_t j = 0; j < K; ++j) key[j] = __shfl(mine, int(g * G + j), kWave);
        // the most requests any of the wave's rows has: the match below
        // compares against that many keys only (most medium rows have one or
        // two deletes, so most waves skip most of the eight compares a slot)
        uint32_t wmax = nd;
        wmax = max(wmax, uint32_t(__shfl_xor(int(wmax), 16, kWave)));
        wmax = max(wmax, uint32_t(__shfl_xor(int(wmax), 32, kWave)));
        uint32_t open = (1u << nd) - 1u;  // uniform within the group
        uint32_t holes = 0;               // uniform within the group
        for (uint32_t q0 = 0; __ballot(open != 0u && q0 < nq) != 0; q0 += G * Q) {
            uint32_t d[Q][4], cand = 0;  // bit 4k+s: slot (k, s) equals an open request
            if (open != 0u && q0 < nq) {
                uint32_t livemask = 0;
#pragma unroll
                for (uint32_t k = 0; k < Q; ++k) {
                    const uint32_t q = q0 + k * G + gl;
                    uint32_t live = load_quad(p, q < nq ? q : nq - 1, deg, d[k]);
                    if (q >= nq) live = 0;
                    livemask |= ((1u << live) - 1u) << (4 * k);
                }
                // claimed and unused requests hold kNone, which no live slot holds
#pragma unroll
                for (uint32_t j = 0; j < K; ++j) {
                    if (j >= wmax) break;  // wave-uniform
#pragma unroll
                    for (uint32_t k = 0; k < Q; ++k)
#pragma unroll
                        for (uint32_t s = 0; s < 4; ++s)
                            if (key[j] == d[k][s]) cand |= 1u << (4 * k + s);
                }
                cand &= livemask;
            }
            // each group claims its candidates one at a time, in lane order
            for (;;) {
                const uint64_t any = __ballot(cand != 0u);
                if (any == 0) break;
                const uint64_t mineg = any & gmask;
                const uint32_t leader = mineg ? uint32_t(__ffsll((unsigned long long)mineg) - 1) : g * G;
                uint32_t claimed = kNone, slot = 0;
                if (mineg && lane == leader) {
                    const uint32_t b = uint32_t(__ffs(cand) - 1);
                    cand &= cand - 1u;
                    const uint32_t x = d[b >> 2][b & 3u];
                    uint32_t m = 0;
#pragma unroll
                    for (uint32_t j = 0; j < K; ++j) m |= (key[j] == x ? 1u : 0u) << j;
                    m &= open;
                    if (m) {
                        claimed = uint32_t(__ffs(m) - 1);
                        slot = 4u * (q0 + (b >> 2) * G + gl) + (b & 3u);
                        p[slot] = kNone;
                    }
                }
                claimed = __shfl(claimed, int(leader), kWave);
                slot = __shfl(slot, int(leader), kWave);
                if (claimed != kNone) {
                    open &= ~(1u << claimed);
#pragma unroll
                    for (uint32_t j = 0; j < K; ++j)
                        if (j == claimed) key[j] = kNone;
                    if (gl == 0) grp[holes] = slot;  // the requests were read into registers above
                    ++holes;
                }
                if (open == 0u) cand = 0;
            }
        }
        if (i < cnt && gl == 0) a.hcnt[u] = holes;  // zero before the call; this group is its only writer
    }
}

template <uint32_t Q, bool HOST, uint32_t K = kDelWaveMax>
__global__ __launch_bounds__(kBlock) void k_del_rows(MutArgs a) { del_rows_body<Q, HOST, K>(a, blockIdx.x, gridDim.x); }

// The tiny-row claims and the host-tier medium rows in one launch (disjoint
// rows): workgroups [0, split) take the records of tiny rows, [split,
// split2) the host-tier rows, whose reads wait on the host link. With M16
// (a large batch) the workgroups from split2 on take the HBM medium rows
// with 9-16 deletes (a launch of their own would add a kernel boundary).
template <bool M16>
__global__ __launch_bounds__(kBlock) void k_del_lane_host(MutArgs a, uint32_t split, uint32_t split2) {
    if (blockIdx.x < split) del_lane_body(a, blockIdx.x, split);
    else if (!M16 || blockIdx.x < split2) del_rows_body<2, true>(a, blockIdx.x - split, split2 - split);
    else if constexpr (M16) del_rows_body<4, false, kDelWaveMax16>(a, blockIdx.x - split2, gridDim.x - split2);
}

// Compaction of rows with at most kDelThreadFix deletes, one thread per row.
// bid / nb: this workgroup's index among the nb workgroups that share the
// work (the whole grid, or its part in k_del_fix_both).
__device__ void del_fix_rows_body(const MutArgs &a, uint32_t bid, uint32_t nb) {
    const uint32_t cnt = ld_acquire(&a.ctl->small_cnt);
    uint64_t applied = 0, missing = 0;
    for (uint32_t t = bid * blockDim.x + threadIdx.x; t < cnt; t += nb * blockDim.x) {
        const uint32_t u = a.small_list[t];
        const uint32_t nd = uint32_t(a.cnt64[u]);
        const uint32_t k = a.hcnt[u];
        if (k) {
            const Row r = a.rows[u];
            uint32_t *row = a.pools.row(r.off);
            const uint32_t *hl = a.grouped + a.gofs[u];
            const uint32_t newdeg = r.deg - k;
            // k <= kDelThreadFix: the k tail slots and the k holes are loaded
            // together first (the stores below go to holes under newdeg, never
            // to the tail), so a host-tier row costs one host-link round trip
            uint32_t tv[kDelThreadFix], hv[kDelThreadFix];
#pragma unroll
            for (uint32_t j = 0; j < kDelThreadFix; ++j) {
                tv[j] = j < k ? row[newdeg + j] : kNone;
                hv[j] = j < k ? hl[j] : kNone;
            }
            // as many holes lie below newdeg as live slots lie at or above
            // it: the j-th live tail slot fills the j-th hole below newdeg
            uint32_t rt = 0;
#pragma unroll
            for (uint32_t j = 0; j < kDelThreadFix; ++j) {
                const bool live = tv[j] != kNone;
                uint32_t target = kNone, rh = 0;
#pragma unroll
                for (uint32_t i = 0; i < kDelThreadFix; ++i) {
                    const bool below = hv[i] < newdeg;
                    if (below && rh == rt) target = hv[i];
                    rh += below ? 1u : 0u;
                }
                if (live && target != kNone) row[target] = tv[j];
                rt += live ? 1u : 0u;
            }
            a.rows[u].deg = newdeg;
            a.hcnt[u] = 0;
        }
        if (u < a.n) {
            applied += k;
            missing += nd - k;
        }
    }
    StatShard *sh = stat_shard(a.stats);
    stat_add(&sh->deletes_applied, wave_sum64(applied));
    stat_add(&sh->missing, wave_sum64(missing));
}

__global__ void k_del_fix_rows(MutArgs a) { del_fix_rows_body(a, blockIdx.x, gridDim.x); }

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

// Long rows, scan phase: one workgroup per kDelTile-slot tile. A slot whose
// destination is in the row's table claims one requested occurrence (first
// claims win). The tile's row loads are issued before the table is staged
// in LDS (tables of more than kDelLdsKeys slots are probed in place).
// An 8192-bit filter of the staged table's keys: a slot whose bit is clear
// matches no request and skips the probe (most slots of a long row match
// nothing, and a wave runs each probe loop as long as its longest chain).
constexpr uint32_t kDelBloomWords = 256;
__device__ __forceinline__ uint32_t del_bloom_bit(uint32_t x) { return (x * 0x85EBCA6Bu) >> 19; }

__global__ __launch_bounds__(256) void k_del_scan(MutArgs a) {
    __shared__ uint32_t keys[kDelLdsKeys];
    // two filters, alternating by tile: a tile fills one while it clears the
    // other for the next tile, so staging needs no barrier of its own
    __shared__ uint32_t blooms[2][kDelBloomWords];
    const uint32_t nt = ld_acquire(&a.ctl->dtile_cnt);
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = tid; k < kDelBloomWords; k += blockDim.x) blooms[0][k] = 0u;
    uint32_t it = 0;
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x, ++it) {
        uint32_t *bloom = blooms[it & 1u], *next = blooms[(it + 1u) & 1u];
        const DelTile dt = a.dtiles[t];
        const uint32_t bits = dt.bits, mask = (1u << bits) - 1u, to = dt.to, len = dt.len;
        uint32_t *row = a.pools.row(dt.off);     // the tile start is a multiple of 4
        const uint32_t nq = (len + 3u) >> 2;     // <= 4 * blockDim.x
        uint32_t d[4][4], live[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t q = k * blockDim.x + tid;
            live[k] = load_quad(row, q < nq ? q : nq - 1, len, d[k]);
            if (q >= nq) live[k] = 0;
        }
        const bool in_lds = mask < kDelLdsKeys;
        __syncthreads();  // the previous tile is done with `keys` and with `next`
        for (uint32_t k = tid; k < kDelBloomWords; k += blockDim.x) next[k] = 0u;
        if (in_lds) {
            for (uint32_t k = tid; k <= mask; k += blockDim.x) {
                const uint32_t key = a.hkeys[to + k];
                keys[k] = key;
                if (key != 0u) {
                    const uint32_t b = del_bloom_bit(key - 1u);
                    atomicOr(&bloom[b >> 5], 1u << (b & 31u));  // cleared by the previous tile
                }
            }
        }
        __syncthreads();
        const uint32_t *gkeys = a.hkeys + to;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
#pragma unroll
            for (uint32_t s = 0; s < 4; ++s) {
                if (s >= live[k]) continue;
                const uint32_t x = d[k][s];
                if (in_lds) {
                    const uint32_t b = del_bloom_bit(x);
                    if (((bloom[b >> 5] >> (b & 31u)) & 1u) == 0u) continue;
                }
                uint32_t h = ht_hash(x, bits), hit = kNone;
                if (in_lds) {
                    for (uint32_t probe = 0; probe <= mask; ++probe, h = (h + 1u) & mask) {
                        const uint32_t key = keys[h];
                        if (key == 0u) break;
                        if (key == x + 1u) {
                            hit = h;
                            break;
                        }
                    }
                } else {
                    for (uint32_t probe = 0; probe <= mask; ++probe, h = (h + 1u) & mask) {
                        const uint32_t key = gkeys[h];
                        if (key == 0u) break;
                        if (key == x + 1u) {
                            hit = h;
                            break;
                        }
                    }
                }
                if (hit != kNone && atomicAdd(&a.hcnts[to + hit].y, 1u) < a.hcnts[to + hit].x) {
                    const uint32_t slot = 4u * (k * blockDim.x + tid) + s;
                    row[slot] = kNone;
                    del_hole(a, dt.row, dt.start + slot);
                }
            }
        }
    }
}

// Compaction of rows with more than kDelThreadFix deletes (or a table), one
// workgroup per row: with k holes the row shrinks to deg-k and the live
// slots of the last k positions move into the holes below deg-k (O(k)
// stores). The row's table and hole counter are cleared for the next call.
__device__ void del_fix_body(const MutArgs &a, uint32_t bid, uint32_t nb) {
    __shared__ uint32_t wsum[256 / kWave];
    const uint32_t cnt = ld_acquire(&a.ctl->big_cnt);
    const uint32_t tid = threadIdx.x;
    uint64_t applied = 0, missing = 0;
    for (uint32_t j = bid; j < cnt; j += nb) {
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
            if (live && run + rk < k) row[hl[run + rk]] = v;
            run += tot;
        }
        // 3. clear the table (long rows) for the next call
        const uint32_t to = a.htoff[u], tsz = to != kNone ? 1u << del_table_bits(nd) : 0u;
        for (uint32_t s = tid; s < tsz; s += blockDim.x) {
            a.hkeys[to + s] = 0u;
            a.hcnts[to + s] = make_uint2(0u, 0u);
        }
        if (tid == 0) {
            a.rows[u].deg = newdeg;
            a.hcnt[u] = 0u;
        }
        if (u < a.n) {
            applied += k;
            missing += nd - k;
        }
        __syncthreads();
    }
    if (wave_in_block() == 0) {
        StatShard *sh = stat_shard(a.stats);
        stat_add(&sh->deletes_applied, applied);
        stat_add(&sh->missing, missing);
    }
}

__global__ __launch_bounds__(256) void k_del_fix(MutArgs a) { del_fix_body(a, blockIdx.x, gridDim.x); }

// Both hole fixes in one launch (they own disjoint rows): the first `split`
// workgroups take the rows fixed by one thread each, the others one row per
// workgroup, so the short per-thread pass and the per-workgroup pass overlap.
__global__ __launch_bounds__(256) void k_del_fix_both(MutArgs a, uint32_t split) {
    if (blockIdx.x < split) del_fix_rows_body(a, blockIdx.x, split);
    else del_fix_body(a, blockIdx.x - split, gridDim.x - split);
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
    uint32_t flags;  // kFlRelease | class: src is a host-tier block to list once copied
};
constexpr uint32_t kFlRelease = 0x80000000u;
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
        // a growing host-tier row takes a freed block of its new class when one is listed
        uint64_t popped = 0;
        const bool reuse = grow && tier == 1 && newcap <= 0xFFFFFFFFull && fl_pop(a.fl, fl_class(newcap), popped);
        const uint64_t o0 = block_reserve<unsigned long long>(grow && tier == 0 ? newcap : 0ull,
                                                              (unsigned long long *)&a.ctl->pool_tail[0], scratch64);
        const uint64_t o1 = block_reserve<unsigned long long>(grow && tier == 1 && !reuse ? newcap : 0ull,
                                                              (unsigned long long *)&a.ctl->pool_tail[1], scratch64);
        const uint64_t o = tier ? (reuse ? popped : o1) : o0;
        const bool over = grow && !reuse && (o + newcap > a.ctl->pool_cap[tier] || newcap > 0xFFFFFFFFull);
        block_append(over, u, a.overflow, &a.ctl->overflow_cnt, scratch);
        const bool go = on && !over && na != 0;
        const uint64_t noff = grow ? (o | (tier ? kHostTier : 0ull)) : r.off;
        const uint32_t np = go ? (grow ? pieces_of(r.deg) : 0u) + pieces_of(na) : 0u;
        const uint32_t pb = block_reserve<uint32_t>(np, &a.ctl->copy_cnt, scratch);
        if (go) {
            const uint32_t *old_row = a.pools.row(r.off);
            uint32_t *row = a.pools.row(noff);
            uint32_t k = pb;
            // the first piece of a relocated host-tier row carries the old
            // block's class: k_fl_release lists the block once it is copied
            if (grow)
                for (uint32_t p = 0; p < r.deg; p += kPieceLen)
                    a.pieces[k++] = AddPiece{old_row + p, row + p, min(kPieceLen, r.deg - p),
                                             p == 0 && tier == 1 ? kFlRelease | fl_class(r.cap) : 0u};
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
                if (u < a.n) acc_reloc += 1;
            } else {
                a.rows[u].deg = uint32_t(newdeg);
            }
            if (u < a.n) acc_adds += na;
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

// Lists the old blocks of the host-tier rows this insert phase relocated
// (their slots were copied by k_add_copy, which ran before).
__global__ void k_fl_release(MutArgs a) {
    const uint32_t cnt = ld_acquire(&a.ctl->copy_cnt);
    for (uint32_t i = gthread(); i < cnt; i += gthreads()) {
        const AddPiece pc = a.pieces[i];
        if (pc.flags & kFlRelease) fl_push(a.fl, pc.flags & 63u, uint64_t(pc.src - a.pools.host));
    }
}

// ------------------------------------------------------- grouped batches
// Several consecutive batches applied as one: every delete is matched to
// the latest earlier unmatched add of the same (u,v) in the group (adds of
// batch b come after its deletes); matched pairs cancel, since the delete
// removes an instance the add created. The records left then give the
// sequential result when all deletes are applied before all adds, so the
// group needs one mutation and one fixed point. Only (u,v) keys that occur
// both as add and as delete in the group are examined one by one.
struct GroupArgs {
    const cig_edge *dels;
    uint32_t nd;
    const cig_edge *adds;
    uint32_t na;
    const uint32_t *del_cut;  // nb+1 prefix of the deletes per batch
    const uint32_t *add_cut;  // nb+1 prefix of the adds per batch
    uint32_t nb;
    uint32_t n;
    unsigned long long *tkey;  // (u<<32 | v) + 1; 0 = empty
    uint32_t *thead;           // op list head (kNone = empty)
    uint8_t *tconf;            // key seen as add and as delete
    uint32_t tbits;
    uint32_t *next_op;         // ops: adds [0,na), deletes [na,na+nd)
    uint8_t *skip;             // 1 = cancelled op
    uint32_t *overflow;        // set when a key holds more than kGroupOps ops
    StatShard *stats;
};
constexpr uint32_t kGroupOps = 32;

__device__ __forceinline__ uint32_t key_hash(unsigned long long k, uint32_t bits) {
    return uint32_t((k * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

__global__ void k_group_adds(GroupArgs g) {
    const uint32_t mask = (1u << g.tbits) - 1u;
    for (uint32_t j = gthread(); j < g.na; j += gthreads()) {
        const cig_edge e = g.adds[j];
        if (e.u >= g.n || e.v >= g.n) continue;
        const unsigned long long key = ((unsigned long long)e.u << 32 | e.v) + 1ull;
        uint32_t h = key_hash(key, g.tbits);
        for (;;) {  // the table has >= 2 * na slots: the probe ends
            const unsigned long long prev = atomicCAS(&g.tkey[h], 0ull, key);
            if (prev == 0ull || prev == key) break;
            h = (h + 1u) & mask;
        }
        g.next_op[j] = atomicExch(&g.thead[h], j);
    }
}

__global__ void k_group_dels(GroupArgs g) {
    const uint32_t mask = (1u << g.tbits) - 1u;
    for (uint32_t i = gthread(); i < g.nd; i += gthreads()) {
        const cig_edge e = g.dels[i];
        if (e.u >= g.n || e.v >= g.n) continue;
        const unsigned long long key = ((unsigned long long)e.u << 32 | e.v) + 1ull;
        for (uint32_t h = key_hash(key, g.tbits), probe = 0; probe <= mask; ++probe, h = (h + 1u) & mask) {
            const unsigned long long k = g.tkey[h];
            if (k == 0ull) break;
            if (k == key) {
                g.tconf[h] = 1;
                g.next_op[g.na + i] = atomicExch(&g.thead[h], g.na + i);
                break;
            }
        }
    }
}

// Batch of entry `i` given the nb+1 prefix `cut` (largest b with cut[b] <= i).
__device__ __forceinline__ uint32_t batch_of(const uint32_t *cut, uint32_t nb, uint32_t i) {
    uint32_t lo = 0, hi = nb;  // cut[lo] <= i < cut[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (cut[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One thread per key seen as both add and delete: order its ops by time
// (delete of batch b at 2b, add at 2b+1) and cancel matched pairs.
__global__ void k_group_match(GroupArgs g) {
    const uint32_t cap = 1u << g.tbits;
    uint64_t cancel_d = 0, cancel_a = 0;
    for (uint32_t h = gthread(); h < cap; h += gthreads()) {
        if (!g.tconf[h]) continue;
        uint32_t op[kGroupOps], tm[kGroupOps], cnt = 0;
        bool over = false;
        for (uint32_t o = g.thead[h]; o != kNone; o = g.next_op[o]) {
            if (cnt == kGroupOps) {
                over = true;
                break;
            }
            const bool del = o >= g.na;
            tm[cnt] = del ? 2u * batch_of(g.del_cut, g.nb, o - g.na) : 2u * batch_of(g.add_cut, g.nb, o) + 1u;
            op[cnt++] = o;
        }
        if (over) {
            atomicExch(g.overflow, 1u);
            continue;
        }
        for (uint32_t x = 1; x < cnt; ++x)  // insertion sort by time
            for (uint32_t y = x; y > 0 && tm[y - 1] > tm[y]; --y) {
                const uint32_t t0 = tm[y]; tm[y] = tm[y - 1]; tm[y - 1] = t0;
                const uint32_t o0 = op[y]; op[y] = op[y - 1]; op[y - 1] = o0;
            }
        uint32_t stack[kGroupOps], top = 0;
        for (uint32_t x = 0; x < cnt; ++x) {
            if (op[x] < g.na) {
                stack[top++] = op[x];
            } else if (top) {
                g.skip[op[x]] = 1;
                g.skip[stack[--top]] = 1;
                ++cancel_d;
                ++cancel_a;
            }
        }
    }
    // a cancelled pair applied both records (the delete found the instance)
    StatShard *sh = stat_shard(g.stats);
    stat_add(&sh->deletes_applied, wave_sum64(cancel_d));
    stat_add(&sh->adds_applied, wave_sum64(cancel_a));
}

// End of a call: clears the per-row counters of the touched rows and, when
// `aff` is given, the affected bitmap words of the affected list. With
// activity counting on, every touched row (outgoing or incoming) adds its
// records of the call to its window count: a delete scans the row, so rows
// mutated often are worth keeping in HBM like rows expanded often.
__global__ void k_clear_touched(MutArgs a, uint32_t *aff, const uint32_t *alist) {
    const uint32_t touched = ld_acquire(&a.ctl->touched);
    for (uint32_t t = gthread(); t < touched; t += gthreads()) {
        const uint32_t u = a.touched[t];
        if (a.hot) {
            const uint64_t c = a.cnt64[u];
            a.hot[u] += uint32_t(min(uint64_t(255), (c & 0xFFFFFFFFull) + (c >> 32)));
        }
        a.cnt64[u] = 0;
    }
    if (aff) {
        const uint32_t cnt = ld_acquire(&a.ctl->acount);
        for (uint32_t i = gthread(); i < cnt; i += gthreads()) aff[alist[i] >> 5] = 0u;
    }
}

// The descriptor and the delete/add counts of each row listed in `list`
// (the rows whose adds found no room), for the host's compaction sizing.
struct RowCount {
    Row row;
    uint64_t cnt;
};
__global__ void k_gather_rows(const Row *rows, const uint64_t *cnt64, const uint32_t *list, uint32_t cnt,
                              RowCount *out) {
    for (uint32_t i = gthread(); i < cnt; i += gthreads()) {
        const uint32_t u = list[i];
        out[i] = RowCount{rows[u], cnt64[u]};
    }
}

// The affected-bit part of k_clear_touched alone.
__global__ void k_clear_aff(const Ctl *c, uint32_t *aff, const uint32_t *alist) {
    const uint32_t cnt = ld_acquire(&c->acount);
    for (uint32_t i = gthread(); i < cnt; i += gthreads()) aff[alist[i] >> 5] = 0u;
}

// ---------------------------------------------------- deletion invalidation
struct InvalArgs {
    Ctl *ctl;
    const Row *rows;
    const Row *in_rows;  // incoming rows: rows + n, or rows itself for a symmetric graph
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
    const uint8_t *add_skip, *del_skip;  // grouped batches: cancelled records (may be null)
    const BatchRef *ref;                 // non-null: the records are read from here (graph replays)
};

// reset_del_edges (driver.cuh:371-411): a deleted (u,v) whose v has parent u
// seeds the affected set. Parallel copies of (u,v) are treated conservatively.
__global__ void k_inval_seed(InvalArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    const uint32_t nd = rec_nd(a);
    const cig_edge *dels = rec_dels(a);
    for (uint32_t base = blockIdx.x * blockDim.x; base < nd; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        bool hit = false;
        uint32_t v = 0;
        if (i < nd) {
            const cig_edge e = dels[i];
            if (e.u < a.n && e.v < a.n && !(a.del_skip && a.del_skip[i]) && uint32_t(a.state[e.v]) == e.u) {
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

// One invalidation wave as an ordinary launch, the seq-th of its chain:
// round k expands its frontier alist[istart, istart+ifront) (the first round:
// the seeds, alist[0, acount)) and its deferred tiles, and appends the next
// frontier right behind it. Slots
// rotate mod 3: workgroup 0 opens slot k+1's start, clears slot k+2 and
// posts the round's progress word at the start, so no workgroup waits for
// another and the round needs no closing step.
template <int QPL>
__global__ __launch_bounds__(kBlock) void k_inval_round(InvalArgs a, uint64_t *prog, uint64_t tag, uint32_t seq) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ RoundStage stage;
    Ctl *c = a.ctl;
    const uint32_t k = c->iround + seq - 1;
    // the first round's frontier is the seeds: alist[0, acount)
    const uint32_t start = seq == 1 ? 0u : c->istart[k % 3];
    const uint32_t cnt = seq == 1 ? c->acount : c->ifront[k % 3];
    const uint32_t nt = c->icount[k % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->istart[(k + 1) % 3] = start + cnt;
        c->ifront[(k + 2) % 3] = 0;
        c->icount[(k + 2) % 3] = 0;
        c->acount = start + cnt;  // affected so far; final once a round finds nothing
        post_progress(prog, tag, cnt != 0 || nt != 0);
    }
    if (cnt == 0 && nt == 0) return;
    InvalidateOp op{a.rows, a.state, a.aff};
    Counters ctr;
    const StagedOut<2048, 128> out{Sink{a.alist + start + cnt, &c->ifront[(k + 1) % 3]},
                                   TileSink{(k & 1) ? a.t0 : a.t1, &c->icount[(k + 1) % 3]}, &stage};
    out_init(out);
    expand_range<QPL>(op, a.cfg, a.alist + start, cnt, out, lds[wave_in_block()], ctr);
    expand_tiles<QPL>(op, a.cfg, (k & 1) ? a.t1 : a.t0, nt, out);
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
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

// SeedAddedEdgeFrontier (framework.cuh:238-294): each inserted (u,v) offers
// value(u)+w to v; improved destinations join the pending frontier.
template <int W, bool kStrict>
__global__ void k_insert_seed(InvalArgs a, uint32_t skip_affected) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    const uint32_t r = a.ctl->round;
    uint32_t *q = (r & 1) ? a.q1 : a.q0;
    const uint32_t na = rec_na(a);
    const cig_edge *adds = rec_adds(a);
    for (uint32_t base = blockIdx.x * blockDim.x; base < na; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        bool push = false;
        uint32_t v = 0;
        if (i < na) {
            const cig_edge e = adds[i];
            if (e.u < a.n && e.v < a.n && !(a.add_skip && a.add_skip[i]) &&
                !(skip_affected && bit_test(a.aff, e.u))) {
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

// -------------------------------------------------------------- pull repair
// With the incoming index on, the repair reads only the incoming rows of the
// affected vertices (GpuAffectedPullRelax, framework.cuh:137-204, over a
// device-resident reverse index instead of the CPU-materialised one of
// dynamic_reverse_index.h:223-274): y takes the minimum over its unaffected,
// reached in-neighbours x of value(x) + w(x, y), ties to the smaller x. This
// is the same candidate set the full scan offers y, so the state after the
// repair is identical.
template <int W, bool kStrict>
struct PullOp {
    const Row *in_rows;  // incoming rows
    uint64_t *state;
    const uint32_t *aff;
    uint64_t seed;
    __device__ bool source(uint32_t y, uint32_t, Row &row, uint32_t &pay) const {
        row = in_rows[y];
        pay = y;
        return row.deg != 0;
    }
    __device__ bool tile_source(const Tile &t, uint32_t &pay) const {
        pay = t.v;
        return true;
    }
    __device__ uint64_t peek(uint32_t x) const { return state[x]; }
    __device__ bool edge(uint32_t y, uint32_t, uint32_t x, uint64_t sx) const {
        const uint32_t vx = uint32_t(sx >> 32);
        if (vx == kInf || bit_test(aff, x)) return false;
        const uint64_t c = uint64_t(vx) + weight<W>(x, y, seed);
        if (c < kInf) relax<kStrict>(state, y, uint32_t(c), x);
        return false;
    }
};

template <int W, bool kStrict, int QPL>
__global__ __launch_bounds__(kBlock) void k_repair_pull(InvalArgs a) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ StageLds<2, 128> stage;
    Ctl *c = a.ctl;
    const uint32_t na = ld_acquire(&c->acount);
    if (na == 0) return;
    PullOp<W, kStrict> op{a.in_rows, a.state, a.aff, a.seed};
    Counters ctr;
    const StagedOut<2, 128> out{Sink{nullptr, &c->scount}, TileSink{a.t0, &c->scount}, &stage};
    out_init(out);
    expand_range<QPL>(op, a.cfg, a.alist, na, out, lds[wave_in_block()], ctr);
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
}

template <int W, bool kStrict, int QPL>
__global__ __launch_bounds__(kBlock) void k_repair_pull_tiles(InvalArgs a) {
    Ctl *c = a.ctl;
    const uint32_t nt = ld_acquire(&c->scount);
    if (nt == 0) return;
    PullOp<W, kStrict> op{a.in_rows, a.state, a.aff, a.seed};
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
    uint32_t delta;  // delta-ordered rounds (0: unordered)
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

// One relax round as an ordinary launch (the default driver), the seq-th of
// its chain: round r = round + seq - 1 consumes queue[r&1] / slot r%3 and
// appends to the other queue / slot (r+1)%3. Workgroup 0 clears slot
// (r+2)%3 and posts the round's progress word at the start; the host keeps
// a few rounds enqueued ahead and advances `round` after the chain, so no
// workgroup waits for another and a round has no closing step. A round
// whose frontier and deferred tiles are empty returns at once.
template <int W, bool kStrict, int QPL, int B = kBlock, bool ORD = false>
__global__ __launch_bounds__(B) void k_relax_round(LoopArgs a, uint64_t *prog, uint64_t tag, uint32_t seq) {
    __shared__ WaveLds lds[B / kWave];
    __shared__ RoundStage stage;
    Ctl *c = a.ctl;
    const uint32_t r = c->round + seq - 1;
    const uint32_t nq = c->qcount[r % 3], nt = c->tcount[r % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->qcount[(r + 2) % 3] = 0;
        c->tcount[(r + 2) % 3] = 0;
        if constexpr (ORD) c->bmin[(r + 1) % 3] = kInf;  // for the next round's k_ord_min
        if (nq || nt) c->rounds += 1;
        post_progress(prog, tag, nq != 0 || nt != 0);
    }
    if (nq == 0 && nt == 0) return;
    Counters ctr;
    RelaxOp<W, kStrict> op{a.rows, a.state, a.stamp, a.hot, a.seed, r + 1};
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    const StagedOut<2048, 128> out{Sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]},
                                   TileSink{(r & 1) ? a.t0 : a.t1, &c->tcount[(r + 1) % 3]}, &stage};
    out_init(out);
    if constexpr (ORD) {
        // Items above the round's smallest bucket wait for a later round
        // (PartitionActive, ordered_gpu_repair.cuh:33-48): re-listed here,
        // skipped by the expansion below. Tiles always expand: their rows
        // were cut in the previous round.
        op.delta = a.delta;
        op.bucket = c->bmin[r % 3];
        const uint32_t wpb = blockDim.x / kWave, gw = blockIdx.x * wpb + wave_in_block();
        for (uint32_t base = gw * kWave; base < nq; base += gridDim.x * wpb * kWave) {
            const uint32_t i = base + lane_id();
            bool f[1] = {false};
            uint32_t v[1] = {0};
            if (i < nq) {
                v[0] = qin[i];
                const uint32_t val = uint32_t(a.state[v[0]] >> 32);
                f[0] = val != kInf && val / a.delta > op.bucket && claim(a.stamp, v[0], r + 1);
            }
            out_push<1>(out, f, v);
        }
    }
    expand_range<QPL>(op, a.cfg, qin, nq, out, lds[wave_in_block()], ctr);
    expand_tiles<QPL>(op, a.cfg, (r & 1) ? a.t1 : a.t0, nt, out);
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
}

// Delta-ordered rounds (FindActiveBucket, ordered_gpu_repair.cuh:28-32):
// the smallest value / delta of round r's frontier, one atomic per wave.
// Values only fall, so a later read in the round can only find the bucket's
// items at or below it: the round always expands at least one row.
__global__ void k_ord_min(LoopArgs a, uint32_t seq) {
    Ctl *c = a.ctl;
    const uint32_t r = c->round + seq - 1;
    const uint32_t nq = c->qcount[r % 3];
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    uint32_t m = kInf;
    for (uint32_t i = gthread(); i < nq; i += gthreads()) {
        const uint32_t val = uint32_t(a.state[qin[i]] >> 32);
        if (val != kInf) m = min(m, val / a.delta);
    }
    for (int off = kWave / 2; off > 0; off >>= 1) m = min(m, uint32_t(__shfl_xor(int(m), off, kWave)));
    if (lane_id() == 0 && m != kInf) atomicMin(&c->bmin[r % 3], m);
}

__global__ void k_ord_reset(Ctl *c) {
    if (threadIdx.x < 3 && blockIdx.x == 0) c->bmin[threadIdx.x] = kInf;
}

// Moves the round counter past a finished chain of `by` rounds.
__global__ void k_advance_round(Ctl *c, uint32_t by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c->round += by;
}

// ---------------------------------------------------------- PageRank rounds
// Pull plan of one solve: incoming rows at least kPullSplit slots long are cut
// into tiles of kPullTile slots; a row of several tiles also gets a PullRow
// whose tiles' partial sums k_pr_pull_long adds up.
constexpr uint32_t kPullSplit = 256;
constexpr uint32_t kPullTile = 2048;

struct PullRow {
    uint32_t v, first, nt, pad;
};

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
    // pull form (incoming rows present): rounds whose frontier holds at
    // least pull_min vertices gather; pull_min = kNone disables it
    const Row *in_rows;
    uint32_t n, pull_min;
    Tile *ptiles;
    PullRow *prow;
    float *tpart;
    // bit v set iff delta[v] != 0; pull passes of rounds with fewer than
    // mask_max frontier vertices (and amendments) test it before gathering
    unsigned long long *amask;
    uint32_t mask_max;
    // out-degrees of the current rows (k_pr_degrees, once per solve): the
    // dense consume reads 4 B per vertex here instead of its 16-B row
    uint32_t *odeg;
};

// Out-degree of every vertex for the dense consume passes of one solve.
__global__ void k_pr_degrees(const Row *rows, uint32_t n, uint32_t *odeg) {
    for (uint32_t v = gthread(); v < n; v += gthreads()) odeg[v] = rows[v].deg;
}

__device__ __forceinline__ bool pr_dense(const PrArgs &a, uint32_t nq) { return nq != 0 && nq >= a.pull_min; }

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
__global__ __launch_bounds__(kBlock) void k_pr_consume(PrArgs a, uint64_t *prog, uint64_t tag, uint32_t seq) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    Ctl *c = a.ctl;
    const uint32_t r = c->round + seq - 1;
    const uint32_t nq = c->qcount[r % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->qcount[(r + 2) % 3] = 0;
        c->tcount[(r + 2) % 3] = 0;
        if (nq) c->rounds += 1;
        post_progress(prog, tag, nq != 0);
    }
    if (nq == 0) return;
    Counters ctr;
    if (pr_dense(a, nq)) {
        // pull form: the queued flags are the frontier; every vertex gets its
        // contribution (0 when not consumed) in delta[v]
        if (blockIdx.x == 0 && threadIdx.x == 0) c->pull_rounds += 1;
        // a wave takes 64 consecutive vertices (one mask word) per step
        for (uint32_t base = gthread() - lane_id(); base < a.n; base += gthreads()) {
            const uint32_t v = base + lane_id();
            float d = 0.0f;
            if (v < a.n && a.queued[v]) {
                a.queued[v] = 0u;
                const float rv = a.res[v];
                if (!isfinite(rv) || !isfinite(a.rank[v])) atomicExch(&c->pr_invalid, 1u);
                if (fabsf(rv) > a.eps) {
                    a.res[v] = 0.0f;
                    const float x = a.rank[v];
                    const float y = rv - a.comp[v];
                    const float t = x + y;
                    a.comp[v] = (t - x) - y;
                    a.rank[v] = t;
                    const uint32_t deg = a.odeg[v];
                    ctr.vexp += 1;
                    if (deg) d = 0.85f * rv / float(deg);
                }
            }
            const unsigned long long bits = __ballot(d != 0.0f);
            if (v < a.n) a.delta[v] = d;
            if (lane_id() == 0) a.amask[base >> 6] = bits;
        }
        flush_counters(ctr, a.cfg.stats, a.cfg.stage);
        return;
    }
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    Tile *tiles = (r & 1) ? a.t1 : a.t0;
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

// PageRank round, scatter phase (pr_residual.cuh:74-100) of round
// round + seq - 1; the consume phase of the same round posted its progress.
template <int QPL>
__global__ __launch_bounds__(kBlock) void k_pr_scatter(PrArgs a, uint32_t seq) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ RoundStage stage;
    Ctl *c = a.ctl;
    const uint32_t r = c->round + seq - 1;
    const uint32_t nq = c->qcount[r % 3];
    if (nq == 0 || pr_dense(a, nq)) return;
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
}

// ------------------------------------------------ PageRank rounds, pull form
// The scatter of a round adds alpha*r(u)/deg(u) to r(w) for every edge (u,w)
// of a consumed u. When the frontier is dense the same sums are gathered
// along the incoming rows (rows n..2n-1, sources of each vertex's in-edges):
// w adds sum over its in-neighbours u of delta[u], which the consume phase
// wrote for every vertex (0 where nothing was consumed). Each r(w) then has
// a single writer, so the residual updates are plain loads and stores, not
// the scatter's per-edge atomics, and a vertex joins the next frontier when
// its residual crosses epsilon, as in the scatter. Only the summation order
// differs from the scatter's (whose atomic order is itself unspecified).

// Builds the pull plan of the current incoming rows (they do not change
// during a solve); `gate` (when given) must be non-zero for it to run.
__global__ void k_pr_pull_plan(PrArgs a, const uint32_t *gate) {
    __shared__ uint32_t scratch[2 * (kMaxWavesPerBlock + 1)];
    Ctl *c = a.ctl;
    if (gate && *gate == 0) return;
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.n; base += gthreads()) {
        const uint32_t v = base + threadIdx.x;
        Row row{};
        if (v < a.n) row = a.in_rows[v];
        const uint32_t nt = row.deg >= kPullSplit ? (row.deg + kPullTile - 1) / kPullTile : 0u;
        const uint32_t x[2] = {nt, nt > 1 ? 1u : 0u};
        uint32_t *const cnt[2] = {&c->pl_tiles, &c->pl_multi};
        uint32_t at[2];
        block_reserve_n<2>(x, cnt, at, scratch);
        const uint32_t mi = nt > 1 ? at[1] : kNone;
        if (nt > 1) a.prow[mi] = PullRow{v, at[0], nt, 0u};
        for (uint32_t j = 0; j < nt; ++j) {
            Tile t;
            t.start = row.off + uint64_t(j) * kPullTile;
            t.v = v;
            t.len = min(kPullTile, row.deg - j * kPullTile);
            t.payload = mi;
            t.pad = 0;
            a.ptiles[at[0] + j] = t;
        }
    }
}

__global__ void k_pr_pull_plan_reset(Ctl *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c->pl_tiles = c->pl_multi = 0;
}

__device__ __forceinline__ float wave_sum_f(float x) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

// Adds `sum` to r(w) (w's only writer this round) and reports whether w
// joins the next frontier.
__device__ __forceinline__ bool pull_apply(const PrArgs &a, uint32_t w, float sum) {
    if (sum == 0.0f) return false;
    const float old = a.res[w];
    const float now = old + sum;
    a.res[w] = now;
    if (!isfinite(now)) atomicExch(&a.ctl->pr_invalid, 1u);
    if (fabsf(old) <= a.eps && fabsf(now) > a.eps && a.queued[w] == 0u) {
        a.queued[w] = 1u;
        return true;
    }
    return false;
}

struct PullLds {
    uint32_t pref[kWave];
    uint32_t deg[kWave];
    uint32_t offlo[kWave];
    uint32_t offhi[kWave];
    float acc[kWave];
};

// Sum of delta[] over the live slots of QPL quads whose loads are in flight
// together; dead slots (past the row's end or past the last quad) add 0.
// With a mask, only sources whose mask bit is set are gathered (the mask is
// 1/32 of delta's size, so a sparse pass touches far fewer cache lines).
template <int QPL>
__device__ __forceinline__ void pull_gather(const float *delta, const unsigned long long *mask,
                                            const uint32_t (&d)[QPL][4], const uint32_t (&live)[QPL],
                                            float (&s)[QPL]) {
    float x[QPL][4];
    if (mask) {
        unsigned long long w[QPL][4];
#pragma unroll
        for (int q = 0; q < QPL; ++q)
#pragma unroll
            for (int k = 0; k < 4; ++k) w[q][k] = mask[d[q][k] >> 6];
#pragma unroll
        for (int q = 0; q < QPL; ++q)
#pragma unroll
            for (int k = 0; k < 4; ++k) x[q][k] = (w[q][k] >> (d[q][k] & 63u)) & 1ull ? delta[d[q][k]] : 0.0f;
    } else {
#pragma unroll
        for (int q = 0; q < QPL; ++q)
#pragma unroll
            for (int k = 0; k < 4; ++k) x[q][k] = delta[d[q][k]];
    }
#pragma unroll
    for (int q = 0; q < QPL; ++q) {
        s[q] = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) s[q] += uint32_t(k) < live[q] ? x[q][k] : 0.0f;
    }
}

// The 64 vertices [v0, v0+64): rows shorter than kPullSplit are summed here,
// lanes walking the wave's concatenated rows quad by quad and adding each
// quad's sum to its row's LDS accumulator.
template <int QPL, class Out>
__device__ void pull_chunk(const PrArgs &a, const unsigned long long *mask, uint32_t v0, PullLds &L,
                           const Out &out, Counters &ctr) {
    const uint32_t lane = lane_id();
    const uint32_t v = v0 + lane;
    uint32_t deg = 0;
    uint64_t off = 0;
    if (v < a.n) {
        const Row row = a.in_rows[v];
        if (row.deg < kPullSplit) {
            deg = row.deg;
            off = row.off;
            ctr.vexp += 1;
        }
    }
    ctr.scanned += deg;
    if (off & kHostTier) ctr.host += deg;
    const uint32_t nq = (deg + 3u) >> 2;
    const uint32_t incl = wave_incl_scan(nq);
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    float sum = 0.0f;
    if (total) {
        L.pref[lane] = incl - nq;
        L.deg[lane] = deg;
        L.offlo[lane] = uint32_t(off);
        L.offhi[lane] = uint32_t(off >> 32);
        L.acc[lane] = 0.0f;
        wave_lds_sync();
        for (uint32_t k0 = 0; k0 < total; k0 += kWave * QPL) {
            uint32_t d[QPL][4], own[QPL], live[QPL];
#pragma unroll
            for (int q = 0; q < QPL; ++q) {
                const uint32_t k = k0 + uint32_t(q) * kWave + lane;
                const uint32_t kk = k < total ? k : total - 1;
                own[q] = owner_of(L.pref, kk);
                const uint64_t o = (uint64_t(L.offhi[own[q]]) << 32) | L.offlo[own[q]];
                live[q] = load_quad_stream(a.cfg.pools.row(o), kk - L.pref[own[q]], L.deg[own[q]], d[q]);
                if (k >= total) live[q] = 0;
            }
            float s[QPL];
            pull_gather<QPL>(a.delta, mask, d, live, s);
#pragma unroll
            for (int q = 0; q < QPL; ++q)
                if (live[q]) atomicAdd(&L.acc[own[q]], s[q]);
        }
        wave_lds_sync();
        sum = L.acc[lane];
        wave_lds_sync();  // LDS is reused by the next chunk of this wave
    }
    const bool f[1] = {deg != 0 && pull_apply(a, v, sum)};
    const uint32_t val[1] = {v};
    out_push<1>(out, f, val);
}

// One tile of a long incoming row: a wave sum; a one-tile row is finished
// here, the tiles of a longer row leave partial sums for k_pr_pull_long.
template <int QPL, class Out>
__device__ void pull_tile(const PrArgs &a, const unsigned long long *mask, uint32_t ti, const Out &out,
                          Counters &ctr) {
    const Tile t = a.ptiles[ti];
    const uint32_t *p = a.cfg.pools.row(t.start);
    const uint32_t lane = lane_id();
    const uint32_t nq = (t.len + 3u) >> 2;
    float acc = 0.0f;
    for (uint32_t k0 = 0; k0 < nq; k0 += kWave * QPL) {
        uint32_t d[QPL][4], live[QPL];
#pragma unroll
        for (int q = 0; q < QPL; ++q) {
            const uint32_t k = k0 + uint32_t(q) * kWave + lane;
            live[q] = load_quad_stream(p, k < nq ? k : nq - 1, t.len, d[q]);
            if (k >= nq) live[q] = 0;
        }
        float s[QPL];
        pull_gather<QPL>(a.delta, mask, d, live, s);
#pragma unroll
        for (int q = 0; q < QPL; ++q) acc += s[q];
    }
    acc = wave_sum_f(acc);
    if (lane == 0) {
        ctr.scanned += t.len;
        if (t.start & kHostTier) ctr.host += t.len;
    }
    bool f[1] = {false};
    if (t.payload == kNone) {
        if (lane == 0) {
            ctr.vexp += 1;
            f[0] = pull_apply(a, t.v, acc);
        }
    } else if (lane == 0) {
        a.tpart[ti] = acc;
    }
    const uint32_t val[1] = {t.v};
    out_push<1>(out, f, val);
}

// Frontier staging of the pull kernels (they make no tiles): small, so that
// many workgroups fit on a CU and keep many gathers in flight.
using PullStage = StageLds<1024, 2>;

// Where a pull pass appends and whether it runs. seq >= 1: the scatter phase
// of round round + seq - 1 when that round is dense, appending to the next
// round's frontier. seq = 0: a batch amendment (gated by amend_pull), whose
// contributions are the touched sources' +-alpha*x/deg, appending to the
// pending frontier of round `round`.
struct PullPass {
    bool run;
    uint32_t stage;
    Sink sink;
    const unsigned long long *mask;
};

__device__ __forceinline__ PullPass pull_pass(const PrArgs &a, uint32_t seq) {
    Ctl *c = a.ctl;
    if (seq == 0) {
        const uint32_t r = c->round;
        return PullPass{c->amend_pull != 0, kStageAmend, Sink{(r & 1) ? a.q1 : a.q0, &c->qcount[r % 3]}, a.amask};
    }
    const uint32_t r = c->round + seq - 1;
    const uint32_t nq = c->qcount[r % 3];
    return PullPass{pr_dense(a, nq), kStageRepair, Sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]},
                    nq < a.mask_max ? a.amask : nullptr};
}

// Pull pass `seq` (see pull_pass): short rows by wave chunks, then the plan's tiles.
template <int QPL>
__global__ __launch_bounds__(kBlock) void k_pr_pull(PrArgs a, uint32_t seq) {
    __shared__ PullLds lds[kWavesPerBlock];
    __shared__ PullStage stage;
    Ctl *c = a.ctl;
    const PullPass pass = pull_pass(a, seq);
    if (!pass.run) return;
    const StagedOut<1024, 2> out{pass.sink, TileSink{nullptr, &c->scount}, &stage};
    out_init(out);
    Counters ctr;
    const uint32_t wpb = blockDim.x / kWave;
    const uint32_t gwave = blockIdx.x * wpb + wave_in_block(), nwaves = gridDim.x * wpb;
    const uint32_t nchunks = (a.n + kWave - 1) / kWave;
    for (uint32_t ch = gwave; ch < nchunks; ch += nwaves)
        pull_chunk<QPL>(a, pass.mask, ch * kWave, lds[wave_in_block()], out, ctr);
    const uint32_t nt = c->pl_tiles;
    for (uint32_t t = gwave; t < nt; t += nwaves) pull_tile<QPL>(a, pass.mask, t, out, ctr);
    flush_counters(ctr, a.cfg.stats, pass.stage);
    out_flush(out);
}

// Rows of several tiles: a wave adds up the tiles' partial sums.
__global__ __launch_bounds__(kBlock) void k_pr_pull_long(PrArgs a, uint32_t seq) {
    __shared__ PullStage stage;
    Ctl *c = a.ctl;
    const uint32_t nm = c->pl_multi;
    const PullPass pass = pull_pass(a, seq);
    if (nm == 0 || !pass.run) return;
    const StagedOut<1024, 2> out{pass.sink, TileSink{nullptr, &c->scount}, &stage};
    out_init(out);
    const uint32_t wpb = blockDim.x / kWave;
    const uint32_t gwave = blockIdx.x * wpb + wave_in_block(), nwaves = gridDim.x * wpb;
    const uint32_t lane = lane_id();
    uint64_t rows_done = 0;
    for (uint32_t i = gwave; i < nm; i += nwaves) {
        const PullRow pr = a.prow[i];
        float s = 0.0f;
        for (uint32_t j = lane; j < pr.nt; j += kWave) s += a.tpart[pr.first + j];
        s = wave_sum_f(s);
        bool f[1] = {false};
        if (lane == 0) {
            rows_done += 1;
            f[0] = pull_apply(a, pr.v, s);
        }
        const uint32_t val[1] = {pr.v};
        out_push<1>(out, f, val);
    }
    Counters ctr;
    ctr.vexp = rows_done;
    flush_counters(ctr, a.cfg.stats, pass.stage);
    out_flush(out);
}

// Batch amendment form (Engine::UpdatePageRank, framework.cuh:1660-1704):
// r += alpha (P_new^T - P_old^T) x over the touched sources, either scattered
// along their old and new rows (few touched rows) or gathered by two sweeps
// of the incoming rows, before and after the mutation (many touched rows: a
// hub-heavy batch touches most edges, and the scatter's atomics would cost
// several sweeps). `touched` counts outgoing and incoming rows.
__global__ void k_pr_amend_form(Ctl *c, uint32_t min_touched) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t pull = c->touched >= min_touched ? 1u : 0u;
    c->amend_pull = pull;
    c->amend_push = 1u - pull;
}

// Pull amendment contributions: delta[u] = sign*alpha*x[u]/deg(u) along u's
// current row for every touched source u, with its mask bit (delta and the
// mask are zero elsewhere).
__global__ void k_pr_amend_delta(PrArgs a, const uint32_t *touched, float sign) {
    Ctl *c = a.ctl;
    if (!c->amend_pull) return;
    const uint32_t cnt = c->touched;
    for (uint32_t i = gthread(); i < cnt; i += gthreads()) {
        const uint32_t u = touched[i];
        if (u >= a.n) continue;  // an incoming row
        const uint32_t deg = a.rows[u].deg;
        const float x = a.rank[u];
        const float d = deg && x != 0.0f ? sign * 0.85f * x / float(deg) : 0.0f;
        a.delta[u] = d;
        if (d != 0.0f) atomicOr(&a.amask[u >> 6], 1ull << (u & 63u));
    }
}

// --------------------------------------------------- multi-source queries
// Up to 64 BFS / SSSP sources at once on the live graph: dist holds 64
// distances per vertex (vertex-major, `stride` words per vertex), act the
// sources whose distance at the vertex improved since it was last expanded.
// A round first takes each frontier vertex's mask (k_ms_consume: snap :=
// act, act := 0), then expands the rows (k_ms_expand): for every edge (u,w)
// and every source s in snap[u], dist[w][s] = min(dist[w][s],
// dist[u][s] + w(u,w)); a destination that improved for any source joins
// the next frontier. Label-correcting per source, so each distance is the
// exact shortest-path value (unit weights: BFS depth).
struct MsArgs {
    Ctl *ctl;
    const Row *rows;
    uint32_t *dist;
    unsigned long long *act, *snap;
    unsigned long long *seen;  // BFS: sources that reached each vertex
    uint32_t *stamp;
    uint32_t *q0, *q1;
    Tile *t0, *t1;
    ExpandCfg cfg;
    uint32_t n, stride;
    uint64_t seed;
};

template <int W>
struct MsOp {
    const Row *rows;
    uint32_t *dist;
    unsigned long long *act;
    const unsigned long long *snap;
    uint32_t *stamp;
    uint32_t stride, tag;
    uint64_t seed;
    __device__ bool source(uint32_t v, uint32_t, Row &row, uint32_t &pay) const {
        if (snap[v] == 0ull) return false;
        row = rows[v];
        pay = 0;
        return row.deg != 0;
    }
    __device__ bool tile_source(const Tile &t, uint32_t &pay) const {
        pay = 0;
        return snap[t.v] != 0ull;
    }
    __device__ uint64_t peek(uint32_t) const { return 0; }
    // Sources are taken eight at a time: their 16 distance loads are in
    // flight together and the minimums are non-returning atomics. A source
    // counts as improved when its candidate beat the value read; if another
    // lane lowered it further meanwhile, w is merely expanded once more for
    // that source with its (smaller) current distance, which changes nothing.
    __device__ bool edge(uint32_t u, uint32_t, uint32_t w, uint64_t) const {
        const uint32_t wt = weight<W>(u, w, seed);
        const uint32_t *du = dist + uint64_t(u) * stride;
        uint32_t *dw = dist + uint64_t(w) * stride;
        unsigned long long gained = 0;
        for (unsigned long long m = snap[u]; m;) {
            constexpr int kB = 8;
            uint32_t s[kB], dv[kB], wv[kB];
            int cnt = 0;
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                s[k] = m ? uint32_t(__ffsll(m) - 1) : 0u;
                cnt += m ? 1 : 0;
                m &= m - 1;
            }
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                dv[k] = k < cnt ? du[s[k]] : kInf;
                wv[k] = k < cnt ? dw[s[k]] : 0u;
            }
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                const uint64_t c = uint64_t(dv[k]) + wt;
                if (k < cnt && c < wv[k]) {  // also rejects c >= kInf
                    __hip_atomic_fetch_min(&dw[s[k]], uint32_t(c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    gained |= 1ull << s[k];
                }
            }
        }
        if (!gained) return false;
        atomicOr(&act[w], gained);
        return claim(stamp, w, tag);
    }
};

// BFS from up to 64 sources, bit-parallel: the rounds are level-synchronous
// (a round expands exactly the bits its vertices gained in the previous
// round, all at the same depth), so a destination's depth for source s is
// fixed the first time bit s reaches it. One 64-bit atomicOr on seen[w] per
// edge finds the newly reached sources; each (w, s) depth is written once by
// the lane that set the bit. Same depths as the per-source relaxation.
struct MsBfsOp {
    const Row *rows;
    uint32_t *dist;
    unsigned long long *act, *seen;
    const unsigned long long *snap;
    uint32_t *stamp;
    uint32_t stride, tag;
    // payload: the depth of the row owner's expanding sources, plus one
    __device__ uint32_t depth_after(uint32_t v) const {
        const uint32_t s = uint32_t(__ffsll(snap[v]) - 1);
        return dist[uint64_t(v) * stride + s] + 1u;
    }
    __device__ bool source(uint32_t v, uint32_t, Row &row, uint32_t &pay) const {
        if (snap[v] == 0ull) return false;
        row = rows[v];
        pay = depth_after(v);
        return row.deg != 0;
    }
    __device__ bool tile_source(const Tile &t, uint32_t &pay) const {
        if (snap[t.v] == 0ull) return false;
        pay = depth_after(t.v);
        return true;
    }
    // seen bits only grow, so a stale read can only under-report them
    __device__ uint64_t peek(uint32_t w) const { return seen[w]; }
    __device__ bool edge(uint32_t u, uint32_t depth, uint32_t w, uint64_t cur) const {
        const unsigned long long bits = snap[u] & ~(unsigned long long)cur;
        if (!bits) return false;
        const unsigned long long gained = bits & ~atomicOr(&seen[w], bits);
        if (!gained) return false;
        uint32_t *dw = dist + uint64_t(w) * stride;
        for (unsigned long long m = gained; m; m &= m - 1) dw[__ffsll(m) - 1] = depth;
        atomicOr(&act[w], gained);
        return claim(stamp, w, tag);
    }
};

template <int W>
struct MsOpOf {
    using type = MsOp<W>;
    __device__ static MsOp<W> make(const MsArgs &a, uint32_t tag) {
        return MsOp<W>{a.rows, a.dist, a.act, a.snap, a.stamp, a.stride, tag, a.seed};
    }
};
template <>
struct MsOpOf<2> {
    using type = MsBfsOp;
    __device__ static MsBfsOp make(const MsArgs &a, uint32_t tag) {
        return MsBfsOp{a.rows, a.dist, a.act, a.seen, a.snap, a.stamp, a.stride, tag};
    }
};

// The sources of a multi-source run, passed by value (256 B of kernel
// arguments), so that no device copy of them is needed.
struct SrcList {
    uint32_t v[64];
};

__global__ void k_ms_seed(MsArgs a, SrcList sources, uint32_t k) {
    const uint32_t r = a.ctl->round;
    for (uint32_t s = gthread(); s < k; s += gthreads()) {
        const uint32_t v = sources.v[s];
        a.dist[uint64_t(v) * a.stride + s] = 0;
        atomicOr(&a.act[v], 1ull << s);
        if (a.seen) atomicOr(&a.seen[v], 1ull << s);
        if (claim(a.stamp, v, r)) {
            const uint32_t at = atomicAdd(&a.ctl->qcount[r % 3], 1u);
            ((r & 1) ? a.q1 : a.q0)[at] = v;
        }
    }
}

// Round round+seq-1, phase 1: snapshot and clear the masks of the frontier.
__global__ void k_ms_consume(MsArgs a, uint64_t *prog, uint64_t tag, uint32_t seq) {
    Ctl *c = a.ctl;
    const uint32_t r = c->round + seq - 1;
    const uint32_t nq = c->qcount[r % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->qcount[(r + 2) % 3] = 0;
        c->scount = 0;
        if (nq) c->rounds += 1;
        post_progress(prog, tag, nq != 0);
    }
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    for (uint32_t i = gthread(); i < nq; i += gthreads()) {
        const uint32_t v = qin[i];
        a.snap[v] = atomicExch(&a.act[v], 0ull);
    }
}

// Phase 2: expand the frontier rows; rows of at least `split` slots become
// tiles for phase 3 of the same round, so every part of a row is expanded
// with the same snapshot.
template <int W, int QPL>
__global__ __launch_bounds__(kBlock) void k_ms_expand(MsArgs a, uint32_t seq) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ RoundStage stage;
    Ctl *c = a.ctl;
    const uint32_t r = c->round + seq - 1;
    const uint32_t nq = c->qcount[r % 3];
    if (nq == 0) return;
    const auto op = MsOpOf<W>::make(a, r + 1);
    Counters ctr;
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    const StagedOut<2048, 128> out{Sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]},
                                   TileSink{a.t0, &c->scount}, &stage};
    out_init(out);
    expand_range<QPL>(op, a.cfg, qin, nq, out, lds[wave_in_block()], ctr);
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
}

// Phase 3: the tiles cut in phase 2.
template <int W, int QPL>
__global__ __launch_bounds__(kBlock) void k_ms_tiles(MsArgs a, uint32_t seq) {
    __shared__ RoundStage stage;
    Ctl *c = a.ctl;
    const uint32_t r = c->round + seq - 1;
    const uint32_t nt = c->scount;
    if (nt == 0) return;
    const auto op = MsOpOf<W>::make(a, r + 1);
    const StagedOut<2048, 128> out{Sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]},
                                   TileSink{nullptr, &c->scount}, &stage};
    out_init(out);
    expand_tiles<QPL>(op, a.cfg, a.t0, nt, out);
    out_flush(out);
}

// Source-major result: out[s * n + v] = dist[v][s].
__global__ void k_ms_gather(const uint32_t *dist, uint32_t n, uint32_t stride, uint32_t k, uint32_t *out) {
    const uint64_t total = uint64_t(n) * k;
    for (uint64_t i = gthread(); i < total; i += gthreads()) {
        const uint32_t s = uint32_t(i / n), v = uint32_t(i % n);
        out[i] = dist[uint64_t(v) * stride + s];
    }
}

// Plain streaming copy (16 B per lane per step): the HBM reference point for
// rocprofv3 FETCH_SIZE / WRITE_SIZE calibration and achievable bandwidth.
__global__ void k_copy_probe(const uint4 *src, uint4 *dst, uint64_t n16) {
    for (uint64_t i = gthread(); i < n16; i += gthreads()) dst[i] = src[i];
}

// Zero-copy read of pinned host memory (16 B per lane per step): the
// host-link reference point for rows in the host tier. The XOR of what each
// thread read goes to `sink` so that the loads stay.
__global__ void k_host_read_probe(const uint4 *src, uint64_t n16, uint32_t *sink) {
    uint32_t x = 0;
    for (uint64_t i = gthread(); i < n16; i += gthreads()) {
        const uint4 v = src[i];
        x ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (x == 0x9E3779B9u) sink[0] = x;  // practically never; keeps the loads
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
        c->touched_sources += v[19];
        c->residual_l1 += l1;
    }
}

// End of a call: the control block goes to mapped pinned host memory, then
// its sequence number (system scope, after a system fence), so the host
// reads it by polling one word instead of a copy and a stream
// synchronisation. One wave; every lane stores its share of the words.
struct CtlPub {
    Ctl c;
    uint64_t seq;
};
__global__ void k_publish_ctl(const Ctl *c, CtlPub *pub, uint64_t seq) {
    static_assert(sizeof(Ctl) % 8 == 0, "whole words");
    constexpr uint32_t kWords = sizeof(Ctl) / 8;
    const uint64_t *src = reinterpret_cast<const uint64_t *>(c);
    uint64_t *dst = reinterpret_cast<uint64_t *>(&pub->c);
    for (uint32_t i = threadIdx.x; i < kWords; i += blockDim.x) dst[i] = src[i];
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(&pub->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
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

// ------------------------------------------------ in-place pool compaction
// Replaces the full-budget compaction copy (driver.cuh:236-247): the rows of
// one tier are bucketed by windows of `wslots` slots of their current offset.
// Window k's rows are packed into a staging buffer and then written to their
// final places, starting at the total capacity of the earlier windows. Rows
// do not overlap, so everything window k writes lies below its last row's
// end, i.e. below every row of window k+1 (not yet moved), and nothing is
// overwritten before it is read. No second pool and no host copy of the rows.
constexpr uint32_t kCmpMaxWin = 1024;
constexpr uint32_t kCmpChunk = 1u << 16;  // rows per workgroup of the bucketing passes

struct CmpJob {
    uint64_t from;   // current offset (tier bit kept)
    uint64_t spos;   // position in the staging buffer = offset within the window's packed range
    uint32_t len;    // live slots
    uint32_t row;
    uint32_t pfirst;  // first of the row's copy pieces
    uint32_t pad;
};

// Rows are copied in pieces of kCmpPiece slots, a wave per piece, so that a
// hub row of millions of slots is spread over the whole chip. Row blocks are
// powers of two of at least 4 slots at offsets that are sums of such blocks,
// so every row (and its staging place) starts 16-B aligned and rounding its
// length up to whole quads stays inside its block.
constexpr uint32_t kCmpPiece = 2048;

struct CmpArgs {
    Row *rows;
    uint32_t nrows, tier, nwin;
    uint64_t wslots;
    unsigned long long *wcnt, *wcap;  // per window: rows, capacity
    unsigned long long *lfill;        // per window: next free list position
    uint32_t *list;                   // row ids grouped by window
    Pools pools;
};

__device__ __forceinline__ bool cmp_in_tier(const Row &r, uint32_t tier) {
    return r.cap != 0 && ((r.off & kHostTier) != 0) == (tier == 1);
}

// Per window: rows and capacity (LDS counts per workgroup chunk).
__global__ void k_cmp_hist(CmpArgs a) {
    __shared__ unsigned long long cnt[kCmpMaxWin], cap[kCmpMaxWin];
    for (uint32_t w = threadIdx.x; w < a.nwin; w += blockDim.x) cnt[w] = cap[w] = 0;
    __syncthreads();
    const uint64_t lo = uint64_t(blockIdx.x) * kCmpChunk, hi = min<uint64_t>(a.nrows, lo + kCmpChunk);
    for (uint64_t v = lo + threadIdx.x; v < hi; v += blockDim.x) {
        const Row r = a.rows[v];
        if (!cmp_in_tier(r, a.tier)) continue;
        const uint32_t w = uint32_t((r.off & ~kHostTier) / a.wslots);
        atomicAdd(&cnt[w], 1ull);
        atomicAdd(&cap[w], (unsigned long long)r.cap);
    }
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < a.nwin; w += blockDim.x) {
        if (cnt[w]) atomicAdd(&a.wcnt[w], cnt[w]);
        if (cap[w]) atomicAdd(&a.wcap[w], cap[w]);
    }
}

// Row ids grouped by window (lfill starts at each window's list base).
__global__ void k_cmp_scatter(CmpArgs a) {
    __shared__ unsigned long long cnt[kCmpMaxWin], base[kCmpMaxWin];
    for (uint32_t w = threadIdx.x; w < a.nwin; w += blockDim.x) cnt[w] = 0;
    __syncthreads();
    const uint64_t lo = uint64_t(blockIdx.x) * kCmpChunk, hi = min<uint64_t>(a.nrows, lo + kCmpChunk);
    for (uint64_t v = lo + threadIdx.x; v < hi; v += blockDim.x) {
        const Row r = a.rows[v];
        if (cmp_in_tier(r, a.tier)) atomicAdd(&cnt[uint32_t((r.off & ~kHostTier) / a.wslots)], 1ull);
    }
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < a.nwin; w += blockDim.x) {
        base[w] = cnt[w] ? atomicAdd(&a.lfill[w], cnt[w]) : 0ull;
        cnt[w] = 0;
    }
    __syncthreads();
    for (uint64_t v = lo + threadIdx.x; v < hi; v += blockDim.x) {
        const Row r = a.rows[v];
        if (!cmp_in_tier(r, a.tier)) continue;
        const uint32_t w = uint32_t((r.off & ~kHostTier) / a.wslots);
        a.list[base[w] + atomicAdd(&cnt[w], 1ull)] = uint32_t(v);
    }
}

// One window: each row's place in the packed range (workgroup-aggregated
// reservations on the window's cursor), its copy pieces and its job.
__global__ void k_cmp_plan(CmpArgs a, const uint32_t *list, uint32_t cnt, unsigned long long *cursor,
                           uint32_t *pcursor, CmpJob *jobs, uint32_t *pjob) {
    __shared__ unsigned long long scratch64[kMaxWavesPerBlock + 1];
    __shared__ uint32_t scratch32[kMaxWavesPerBlock + 1];
    for (uint32_t base = blockIdx.x * blockDim.x; base < cnt; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        Row r{};
        uint32_t v = 0, np = 0;
        if (i < cnt) {
            v = list[i];
            r = a.rows[v];
            np = max(1u, (r.deg + kCmpPiece - 1) / kCmpPiece);  // an empty row still gets its offset
        }
        const uint64_t pos = block_reserve<unsigned long long>(i < cnt ? r.cap : 0ull, cursor, scratch64);
        const uint32_t pf = block_reserve<uint32_t>(np, pcursor, scratch32);
        if (i < cnt) {
            jobs[i] = CmpJob{r.off, pos, r.deg, v, pf, 0u};
            for (uint32_t k = 0; k < np; ++k) pjob[pf + k] = i;
        }
    }
}

// Copies slots [k*kCmpPiece, ...) of a row, whole quads, by one wave.
__device__ __forceinline__ void cmp_copy_piece(uint32_t *d, const uint32_t *s, uint32_t len, uint32_t k) {
    const uint32_t lo = k * kCmpPiece, hi = min(len, lo + kCmpPiece);
    if (hi <= lo) return;
    const uint32_t nq = (hi - lo + 3u) >> 2;
    const uint4 *sq = reinterpret_cast<const uint4 *>(s + lo);
    uint4 *dq = reinterpret_cast<uint4 *>(d + lo);
    for (uint32_t q = lane_id(); q < nq; q += kWave) dq[q] = sq[q];
}

// Live slots into the staging buffer (a wave per piece).
__global__ void k_cmp_in(CmpArgs a, const CmpJob *jobs, const uint32_t *pjob, const uint32_t *pcount,
                         uint32_t *stage) {
    const uint32_t np = *pcount;
    for (uint32_t p = gwave(); p < np; p += gwaves()) {
        const CmpJob jb = jobs[pjob[p]];
        cmp_copy_piece(stage + jb.spos, a.pools.row(jb.from), jb.len, p - jb.pfirst);
    }
}

// From the staging buffer to the final place, which the row now names.
__global__ void k_cmp_out(CmpArgs a, const CmpJob *jobs, const uint32_t *pjob, const uint32_t *pcount,
                          const uint32_t *stage, uint64_t newbase) {
    const uint64_t tbit = a.tier ? kHostTier : 0ull;
    const uint32_t np = *pcount;
    for (uint32_t p = gwave(); p < np; p += gwaves()) {
        const CmpJob jb = jobs[pjob[p]];
        const uint64_t off = (newbase + jb.spos) | tbit;
        cmp_copy_piece(a.pools.row(off), stage + jb.spos, jb.len, p - jb.pfirst);
        if (p == jb.pfirst && lane_id() == 0) a.rows[jb.row].off = off;
    }
}

// Lists the host-tier blocks that rows loaded into HBM vacated (k_move_rows,
// which copied them, ran before).
__global__ void k_move_release(const CopyJob *jobs, const uint32_t *njobs, FreeList fl) {
    const uint32_t cnt = ld_acquire(njobs);
    for (uint32_t j = gthread(); j < cnt; j += gthreads()) {
        const CopyJob jb = jobs[j];
        if ((jb.from & kHostTier) && jb.cap) fl_push(fl, fl_class(jb.cap), jb.from & ~kHostTier);
    }
}

// Moves rows between tiers (cache admission / eviction). Each job names a row
// whose descriptor has already been rewritten; `from` is the old location.
// A wave per copy piece (see kCmpPiece), whole quads.
__global__ void k_move_rows(const CopyJob *jobs, const uint32_t *pjob, const uint32_t *pcount, Pools pools) {
    const uint32_t np = ld_acquire(pcount);
    for (uint32_t p = gwave(); p < np; p += gwaves()) {
        const CopyJob jb = jobs[pjob[p]];
        cmp_copy_piece(pools.row(jb.to), pools.row(jb.from), jb.len, p - jb.pfirst);
    }
}

// ------------------------------------------------------- cache admission
// Device form of compute_hot_vertices + confirm_candidate_batch
// (framework.cuh:605-729): rows are admitted to HBM by decreasing activity
// score until 90% of the budget is used. Scores are sums of four 8-bit
// window counts (<= 1020), so a per-score capacity histogram replaces the
// reference's sort: the host walks it from the top to find the threshold
// score, rows above it are admitted, rows at it share what room is left.
constexpr uint32_t kScoreBuckets = 1024;

__global__ void k_adm_hist(const Row *rows, const uint32_t *score, uint32_t nrows, unsigned long long *hist) {
    __shared__ unsigned long long h[kScoreBuckets];
    for (uint32_t i = threadIdx.x; i < kScoreBuckets; i += blockDim.x) h[i] = 0ull;
    __syncthreads();
    for (uint32_t v = gthread(); v < nrows; v += gthreads()) {
        const uint32_t cap = rows[v].cap;
        if (cap) atomicAdd(&h[min(score[v], kScoreBuckets - 1)], (unsigned long long)cap);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kScoreBuckets; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

struct AdmArgs {
    const Row *rows;
    const uint32_t *score;
    uint32_t nrows;
    uint32_t thr;              // rows scoring above thr are admitted, below it not
    unsigned long long room;   // slots for rows scoring exactly thr
    unsigned long long *acc;   // [0] room used, [1] evict slots, [2] load slots
    uint32_t *evict, *load;    // rows leaving / entering HBM
    uint32_t *evict_cnt, *load_cnt;
};

// Lists the rows that change tier. Rows scoring exactly thr take the room
// in two passes, rows already in HBM first (pass 0), so that a threshold
// bucket that fits moves nothing; the room is reserved per workgroup step.
__global__ void k_adm_mark(AdmArgs a, uint32_t pass) {
    __shared__ unsigned long long s64[kMaxWavesPerBlock + 1];
    __shared__ uint32_t s32[kMaxWavesPerBlock + 1];
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.nrows; base += gthreads()) {
        const uint32_t v = base + threadIdx.x;
        Row r{};
        uint32_t sc = 0;
        if (v < a.nrows) {
            r = a.rows[v];
            sc = min(a.score[v], kScoreBuckets - 1);
        }
        const bool cached = r.cap && !(r.off & kHostTier);
        const bool cand = r.cap && sc == a.thr && (pass == 0 ? cached : !cached);
        const unsigned long long at =
            block_reserve<unsigned long long>(cand ? (unsigned long long)r.cap : 0ull, &a.acc[0], s64);
        bool decided = false, admit = false;
        if (r.cap && sc != a.thr && pass == 0) {
            decided = true;
            admit = sc > a.thr;
        } else if (cand) {
            decided = true;
            admit = at + r.cap <= a.room;
        }
        const bool ev = decided && cached && !admit, ld = decided && !cached && admit;
        block_reserve<unsigned long long>(ev ? (unsigned long long)r.cap : 0ull, &a.acc[1], s64);
        block_reserve<unsigned long long>(ld ? (unsigned long long)r.cap : 0ull, &a.acc[2], s64);
        block_append(ev, v, a.evict, a.evict_cnt, s32);
        block_append(ld, v, a.load, a.load_cnt, s32);
    }
}

// New blocks in `tier` for the listed rows (bump allocation from the tier's
// tail, one reservation per workgroup step), the copy jobs, the rows' new
// offsets; the vacated blocks become garbage of the other tier.
__global__ void k_move_plan(Row *rows, const uint32_t *list, const uint32_t *count, uint32_t tier, Ctl *c,
                            CopyJob *jobs, uint32_t *pcount, uint32_t *pjob, FreeList fl) {
    __shared__ unsigned long long s64[kMaxWavesPerBlock + 1];
    __shared__ uint32_t s32[kMaxWavesPerBlock + 1];
    const uint32_t cnt = ld_acquire(count);
    for (uint32_t base = blockIdx.x * blockDim.x; base < cnt; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        Row r{};
        uint32_t v = 0, np = 0;
        uint64_t popped = 0;
        bool reuse = false;
        if (i < cnt) {
            v = list[i];
            r = rows[v];
            np = max(1u, (r.deg + kCmpPiece - 1) / kCmpPiece);
            // an evicted row takes a freed host-tier block of its class when one is listed
            reuse = tier == 1 && fl_pop(fl, fl_class(r.cap), popped);
        }
        const unsigned long long bumped = block_reserve<unsigned long long>(
            reuse ? 0ull : (unsigned long long)r.cap, reinterpret_cast<unsigned long long *>(&c->pool_tail[tier]),
            s64);
        const unsigned long long at = reuse ? popped : bumped;
        block_reserve<unsigned long long>((unsigned long long)r.cap,
                                          reinterpret_cast<unsigned long long *>(&c->garbage[1 - tier]), s64);
        const uint32_t pf = block_reserve<uint32_t>(np, pcount, s32);
        if (i < cnt) {
            CopyJob j;
            j.from = r.off;
            j.to = at | (tier ? kHostTier : 0ull);
            j.len = r.deg;
            j.pfirst = pf;
            j.cap = r.cap;
            j.pad = 0;
            jobs[i] = j;
            rows[v].off = j.to;
            for (uint32_t k = 0; k < np; ++k) pjob[pf + k] = i;
        }
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
