// Device building blocks of the incremental graph engine (gfx950, wave64).
//
// Frontier expansion is wave-cooperative: a wave takes 64 frontier items,
// scans their row lengths in registers, stages (prefix, owner, payload, row
// address) in LDS and then walks the concatenated rows 64*EPL edges at a time,
// so consecutive lanes read consecutive 4-B slots of a row (256-B coalesced
// wave reads). Rows at least `split` long are cut into tiles of `tile_len`
// slots that any wave of the grid expands. Outputs are appended with one
// atomic per wave (ballot + popcount ranks). This replaces the reference's
// CTA/warp degree scheduler (include/groute/device/cta_scheduler_hybrid.cuh:
// 59-229) and its per-item atomicAdd queue append (queue.cuh:84-88).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace cig {
namespace dev {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t wave_in_block() { return threadIdx.x / kWave; }

template <class T>
__device__ __forceinline__ T ld_acquire(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, kWave);
        if (lane >= uint32_t(d)) x += y;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t x) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

// Orders this wave's LDS stores before its following LDS loads.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Appends `val` for every lane whose flag is set: one atomic per wave.
// Must be reached by all lanes of the wave (uniform control flow).
__device__ __forceinline__ void wave_append(bool flag, uint32_t val, uint32_t *list,
                                            uint32_t *counter) {
    const uint64_t mask = __ballot(flag);
    if (mask == 0) return;
    const uint32_t leader = uint32_t(__ffsll((unsigned long long)mask) - 1);
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
    base = __shfl(base, int(leader), kWave);
    if (flag) list[base + uint32_t(__popcll(mask & __lanemask_lt()))] = val;
}

struct Pools {
    uint32_t *hbm;
    uint32_t *host;  // device-visible pointer of the pinned host tier
    __device__ __forceinline__ uint32_t *row(uint64_t off) const {
        return (off & kHostTier) ? host + (off & ~kHostTier) : hbm + off;
    }
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Edge weight kinds: 0 = (u+v)%128+1 (reference TraversalEdgeWeight,
// api.cuh:26-28), 1 = 1+splitmix64(seed^(u<<32|v))%64, 2 = unit (BFS),
// 3 = zero (CC). 2^32 is a multiple of 128, so the 32-bit wrapped sum has the
// same low 7 bits as the 64-bit sum the reference's checker uses.
template <int W>
__device__ __forceinline__ uint32_t weight(uint32_t u, uint32_t v, uint64_t seed) {
    if constexpr (W == 0) return ((u + v) & 127u) + 1u;
    else if constexpr (W == 1) return 1u + uint32_t(splitmix64(seed ^ (uint64_t(u) << 32 | v)) & 63u);
    else if constexpr (W == 2) return 1u;
    else return 0u;
}

// State word: value (distance / depth / label) in the high half, parent in the
// low half. kStrict=false: one 64-bit atomicMin; ties on the value keep the
// smaller parent, which stays acyclic because weights are >= 1. kStrict=true
// (zero weights, CC): the pair is replaced only on a strict value decrease,
// so a parent always got its value earlier than its child (acyclic tree).
template <bool kStrict>
__device__ __forceinline__ bool relax(uint64_t *state, uint32_t dst, uint32_t cand,
                                      uint32_t src) {
    uint64_t cur = state[dst];  // may be stale-high inside a fixed point: safe
    if (uint32_t(cur >> 32) <= cand) return false;
    const uint64_t nv = (uint64_t(cand) << 32) | src;
    if constexpr (!kStrict) {
        const uint64_t old = atomicMin((unsigned long long *)&state[dst], (unsigned long long)nv);
        return uint32_t(old >> 32) > cand;
    } else {
        for (;;) {
            const uint64_t prev = atomicCAS((unsigned long long *)&state[dst],
                                            (unsigned long long)cur, (unsigned long long)nv);
            if (prev == cur) return true;
            cur = prev;
            if (uint32_t(cur >> 32) <= cand) return false;
        }
    }
}

// Frontier dedup: true for the first claimant of `tag` on v.
__device__ __forceinline__ bool claim(uint32_t *stamp, uint32_t v, uint32_t tag) {
    if (stamp[v] == tag) return false;
    return atomicExch(&stamp[v], tag) != tag;
}

__device__ __forceinline__ bool bit_test(const uint32_t *bits, uint32_t v) {
    return (bits[v >> 5] >> (v & 31)) & 1u;
}
__device__ __forceinline__ bool bit_claim(uint32_t *bits, uint32_t v) {
    const uint32_t m = 1u << (v & 31);
    if (bits[v >> 5] & m) return false;
    return !(atomicOr(&bits[v >> 5], m) & m);
}

// ---------------------------------------------------------------- operators
// source(): decide whether item v (at list position i) expands; fill its row
// and a 32-bit payload. tile_source(): payload for a tile of v's row.
// peek(): the per-edge load the decision needs (issued for all edges a lane
// holds before any decision, so the loads overlap). edge(): per-edge work on
// the peeked word; returns true to append the destination.

// Push relax (SyncPushDDB relax body + AccumulateBuffer,
// sync_push_dd.cuh:148-239, hybrid_sssp.cu:108-128).
template <int W, bool kStrict>
struct RelaxOp {
    const Row *rows;
    uint64_t *state;
    uint32_t *stamp;
    uint32_t *hot;
    uint64_t seed;
    uint32_t tag;
    __device__ bool source(uint32_t v, uint32_t, Row &row, uint32_t &pay) const {
        const uint32_t val = uint32_t(state[v] >> 32);
        if (val == kInf) return false;
        row = rows[v];
        pay = val;
        if (hot) hot[v] += 1;
        return row.deg != 0;
    }
    __device__ bool tile_source(const Tile &t, uint32_t &pay) const {
        pay = uint32_t(state[t.v] >> 32);
        return pay != kInf;
    }
    __device__ uint64_t peek(uint32_t d) const { return state[d]; }
    __device__ bool edge(uint32_t u, uint32_t val, uint32_t d, uint64_t cur) const {
        const uint64_t c = uint64_t(val) + weight<W>(u, d, seed);
        if (c >= uint32_t(cur >> 32)) return false;  // also rejects c >= kInf
        if (!relax<kStrict>(state, d, uint32_t(c), u)) return false;
        return claim(stamp, d, tag);
    }
};

// Deletion invalidation wave (RelaxCTADB_del + PushFunctorDB_Del,
// sync_push_dd.cuh:1220-1303, push_functor.h:1106-1140): a dependent
// destination is one whose parent is the expanding vertex.
struct InvalidateOp {
    const Row *rows;
    const uint64_t *state;
    uint32_t *aff;  // bitmap
    __device__ bool source(uint32_t v, uint32_t, Row &row, uint32_t &pay) const {
        row = rows[v];
        pay = 0;
        return row.deg != 0;
    }
    __device__ bool tile_source(const Tile &, uint32_t &pay) const {
        pay = 0;
        return true;
    }
    __device__ uint64_t peek(uint32_t y) const { return state[y]; }
    __device__ bool edge(uint32_t x, uint32_t, uint32_t y, uint64_t cur) const {
        if (uint32_t(cur) != x) return false;
        return bit_claim(aff, y);
    }
};

// PageRank residual scatter (pr_residual.cuh:21-30, 74-100): add the
// contribution and enqueue a vertex whose |r| crosses epsilon.
struct PrScatterOp {
    const Row *rows;
    float *res;
    uint32_t *queued;
    const float *delta;  // per frontier position (consume phase output)
    uint32_t *invalid;
    float eps;
    __device__ bool source(uint32_t v, uint32_t i, Row &row, uint32_t &pay) const {
        const float d = delta[i];
        if (d == 0.0f) return false;
        row = rows[v];
        pay = __float_as_uint(d);
        return row.deg != 0;
    }
    __device__ bool tile_source(const Tile &t, uint32_t &pay) const {
        pay = t.payload;
        return true;
    }
    __device__ uint64_t peek(uint32_t) const { return 0; }
    __device__ bool edge(uint32_t, uint32_t pay, uint32_t w, uint64_t) const {
        const float c = __uint_as_float(pay);
        const float old = unsafeAtomicAdd(&res[w], c);
        const float now = old + c;
        if (!isfinite(now)) atomicExch(invalid, 1u);
        if (fabsf(old) <= eps && fabsf(now) > eps) return atomicCAS(&queued[w], 0u, 1u) == 0u;
        return false;
    }
};

// PageRank amendment of touched sources (Runtime::Amend, pr_residual.cuh:160-166):
// +/- alpha * x[u] / deg(u) along u's row.
struct PrAmendOp {
    const Row *rows;
    const float *rank;
    float sign;
    PrScatterOp scatter;
    __device__ bool source(uint32_t v, uint32_t, Row &row, uint32_t &pay) const {
        row = rows[v];
        const float x = rank[v];
        if (row.deg == 0 || x == 0.0f) return false;
        pay = __float_as_uint(sign * 0.85f * x / float(row.deg));
        return true;
    }
    __device__ bool tile_source(const Tile &t, uint32_t &pay) const {
        pay = t.payload;
        return true;
    }
    __device__ uint64_t peek(uint32_t) const { return 0; }
    __device__ bool edge(uint32_t u, uint32_t pay, uint32_t w, uint64_t p) const {
        return scatter.edge(u, pay, w, p);
    }
};

// ---------------------------------------------------------------- expansion
struct WaveLds {
    uint32_t pref[kWave];
    uint32_t v[kWave];
    uint32_t pay[kWave];
    uint32_t offlo[kWave];
    uint32_t offhi[kWave];
};

struct ExpandCfg {
    Pools pools;
    uint32_t split;     // rows with deg >= split go to tiles
    uint32_t tile_len;  // slots per tile
    uint32_t stage;     // Stage whose counters this pass feeds
};

struct Sink {
    uint32_t *list;
    uint32_t *count;
};

struct TileSink {
    Tile *tiles;
    uint32_t *count;
};

struct Counters {
    uint64_t scanned = 0, host = 0, vexp = 0;
};

// Largest j with pref[j] <= k (pref is a non-decreasing exclusive scan with
// pref[0] = 0 <= k); entries of empty rows share their successor's prefix,
// so j always owns a non-empty row.
__device__ __forceinline__ uint32_t owner_of(const uint32_t *pref, uint32_t k) {
    // j only grows by 32+16+8+4+2+1 = 63, so every probe pref[j+step] is in range
    uint32_t j = 0;
#pragma unroll
    for (uint32_t step = 32; step > 0; step >>= 1)
        if (pref[j + step] <= k) j += step;
    return j;
}

// Expands the 64 items [base, base+64) of `items` (or the identity range when
// items == nullptr) that are below `count`.
template <int EPL, class Op>
__device__ void expand_chunk(const Op &op, const ExpandCfg &cfg, const uint32_t *items,
                             uint32_t base, uint32_t count, WaveLds &L, const Sink &sink,
                             const TileSink &tsink, Counters &ctr) {
    const uint32_t lane = lane_id();
    const uint32_t i = base + lane;
    uint32_t v = 0, pay = 0, deg = 0;
    uint64_t off = 0;
    if (i < count) {
        v = items ? items[i] : i;
        Row row;
        if (op.source(v, i, row, pay)) {
            deg = row.deg;
            off = row.off;
        }
    }
    ctr.scanned += deg;
    ctr.vexp += deg ? 1 : 0;
    if (off & kHostTier) ctr.host += deg;
    // long rows become tiles; one tile-count atomic per wave
    const uint32_t nt = deg >= cfg.split ? (deg + cfg.tile_len - 1) / cfg.tile_len : 0u;
    const uint32_t tincl = wave_incl_scan(nt);
    const uint32_t ttotal = __shfl(tincl, kWave - 1, kWave);
    uint32_t tbase = 0;
    if (ttotal && lane == 0) tbase = atomicAdd(tsink.count, ttotal);
    tbase = __shfl(tbase, 0, kWave);
    if (nt) {
        const uint32_t t0 = tbase + tincl - nt;
        for (uint32_t j = 0; j < nt; ++j) {
            Tile t;
            t.start = off + uint64_t(j) * cfg.tile_len;
            t.v = v;
            t.len = min(cfg.tile_len, deg - j * cfg.tile_len);
            t.payload = pay;
            t.pad = 0;
            tsink.tiles[t0 + j] = t;
        }
        deg = 0;
    }
    const uint32_t incl = wave_incl_scan(deg);
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    if (total == 0) return;
    L.pref[lane] = incl - deg;
    L.v[lane] = v;
    L.pay[lane] = pay;
    L.offlo[lane] = uint32_t(off);
    L.offhi[lane] = uint32_t(off >> 32);
    wave_lds_sync();
    for (uint32_t k0 = 0; k0 < total; k0 += kWave * EPL) {
        uint32_t dst[EPL], own[EPL];
        uint64_t cur[EPL];
        bool ok[EPL];
        // owner searches and loads are unconditional (indices clamped to the
        // last edge) so that all EPL chains are in flight together
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
            const uint32_t k = k0 + uint32_t(e) * kWave + lane;
            ok[e] = k < total;
            const uint32_t kk = ok[e] ? k : total - 1;
            own[e] = owner_of(L.pref, kk);
            const uint64_t o = (uint64_t(L.offhi[own[e]]) << 32) | L.offlo[own[e]];
            dst[e] = cfg.pools.row(o)[kk - L.pref[own[e]]];
        }
#pragma unroll
        for (int e = 0; e < EPL; ++e) cur[e] = op.peek(dst[e]);
        bool push[EPL];
#pragma unroll
        for (int e = 0; e < EPL; ++e)
            push[e] = ok[e] && op.edge(L.v[own[e]], L.pay[own[e]], dst[e], cur[e]);
#pragma unroll
        for (int e = 0; e < EPL; ++e) wave_append(push[e], dst[e], sink.list, sink.count);
    }
    wave_lds_sync();  // LDS is reused by the next chunk of this wave
}

template <int EPL, class Op>
__device__ void expand_tile(const Op &op, const ExpandCfg &cfg, const Tile &t, const Sink &sink) {
    uint32_t pay;
    if (!op.tile_source(t, pay)) return;
    const uint32_t *p = cfg.pools.row(t.start);
    const uint32_t lane = lane_id();
    for (uint32_t k0 = 0; k0 < t.len; k0 += kWave * EPL) {
        uint32_t dst[EPL];
        uint64_t cur[EPL];
        bool ok[EPL];
#pragma unroll
        for (int e = 0; e < EPL; ++e) {
            const uint32_t k = k0 + uint32_t(e) * kWave + lane;
            ok[e] = k < t.len;
            dst[e] = p[ok[e] ? k : t.len - 1];
        }
#pragma unroll
        for (int e = 0; e < EPL; ++e) cur[e] = op.peek(dst[e]);
        bool push[EPL];
#pragma unroll
        for (int e = 0; e < EPL; ++e) push[e] = ok[e] && op.edge(t.v, pay, dst[e], cur[e]);
#pragma unroll
        for (int e = 0; e < EPL; ++e) wave_append(push[e], dst[e], sink.list, sink.count);
    }
}

// Grid-strided driver over a list (or identity range) and then a tile list.
template <int EPL, class Op>
__device__ void expand_range(const Op &op, const ExpandCfg &cfg, const uint32_t *items,
                             uint32_t count, const Sink &sink, const TileSink &tsink,
                             WaveLds &L, Counters &ctr) {
    const uint32_t gwave = blockIdx.x * kWavesPerBlock + wave_in_block();
    const uint32_t nwaves = gridDim.x * kWavesPerBlock;
    for (uint64_t base = uint64_t(gwave) * kWave; base < count; base += uint64_t(nwaves) * kWave)
        expand_chunk<EPL>(op, cfg, items, uint32_t(base), count, L, sink, tsink, ctr);
}

template <int EPL, class Op>
__device__ void expand_tiles(const Op &op, const ExpandCfg &cfg, const Tile *tiles,
                             uint32_t ntiles, const Sink &sink, Counters &ctr) {
    const uint32_t gwave = blockIdx.x * kWavesPerBlock + wave_in_block();
    const uint32_t nwaves = gridDim.x * kWavesPerBlock;
    for (uint32_t t = gwave; t < ntiles; t += nwaves) {
        const Tile tile = tiles[t];
        expand_tile<EPL>(op, cfg, tile, sink);
    }
    (void)ctr;
}

// Adds this wave's counters into the control block (one atomic per wave and
// counter), both to the call totals and to the pass's stage.
__device__ __forceinline__ void flush_counters(Counters &ctr, Ctl *ctl, uint32_t stage) {
    const uint64_t s = wave_sum64(ctr.scanned), h = wave_sum64(ctr.host), x = wave_sum64(ctr.vexp);
    if (lane_id() == 0) {
        if (s) {
            atomicAdd((unsigned long long *)&ctl->edges_scanned, (unsigned long long)s);
            atomicAdd((unsigned long long *)&ctl->st_edges[stage], (unsigned long long)s);
        }
        if (h) {
            atomicAdd((unsigned long long *)&ctl->host_edges, (unsigned long long)h);
            atomicAdd((unsigned long long *)&ctl->st_host[stage], (unsigned long long)h);
        }
        if (x) {
            atomicAdd((unsigned long long *)&ctl->vexp, (unsigned long long)x);
            atomicAdd((unsigned long long *)&ctl->st_vexp[stage], (unsigned long long)x);
        }
    }
    ctr = Counters{};
}

}  // namespace dev
}  // namespace cig
