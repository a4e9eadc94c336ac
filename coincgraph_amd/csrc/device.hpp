// Device building blocks of the incremental graph engine (gfx950, wave64).
//
// Frontier expansion is wave-cooperative: a wave takes 64 frontier items,
// scans their row lengths (in 4-slot quads) in registers, stages (prefix,
// owner, payload, degree, row address) in LDS and then walks the concatenated
// rows 64*QPL quads at a time, so consecutive lanes read consecutive 16-B
// quads of a row (1-KB coalesced wave reads, one owner search per quad). Rows at least `split` long are cut into tiles of `tile_len`
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
__device__ __forceinline__ bool relax_from(uint64_t *state, uint32_t dst, uint32_t cand,
                                           uint32_t src, uint64_t cur) {
    // cur: a value of state[dst] read earlier (may be stale-high inside a
    // fixed point: safe, the atomic decides)
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

template <bool kStrict>
__device__ __forceinline__ bool relax(uint64_t *state, uint32_t dst, uint32_t cand,
                                      uint32_t src) {
    return relax_from<kStrict>(state, dst, cand, src, state[dst]);
}

// Frontier dedup: true for the first claimant of `tag` on v.
__device__ __forceinline__ bool claim(uint32_t *stamp, uint32_t v, uint32_t tag) {
    if (stamp[v] == tag) return false;
    return atomicExch(&stamp[v], tag) != tag;
}
// claim() without the test load, for a thread that has just improved v: the
// first claimant of a round usually finds the tag absent, so the exchange
// alone saves a dependent memory round trip.
__device__ __forceinline__ bool claim_now(uint32_t *stamp, uint32_t v, uint32_t tag) {
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
// bit_claim() without the test load, where the bit is usually still clear.
__device__ __forceinline__ bool bit_claim_now(uint32_t *bits, uint32_t v) {
    const uint32_t m = 1u << (v & 31);
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
    // Delta-ordered rounds: only rows whose value / delta is at most `bucket`
    // expand; the rest were deferred to the next round (delta 0: unordered).
    uint32_t delta = 0, bucket = kInf;
    __device__ bool source(uint32_t v, uint32_t, Row &row, uint32_t &pay) const {
        // both loads issued before either is used (one memory round trip)
        const Row r = rows[v];
        const uint32_t val = uint32_t(state[v] >> 32);
        if (val == kInf) return false;
        if (delta && val / delta > bucket) return false;
        row = r;
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
        if (!relax_from<kStrict>(state, d, uint32_t(c), u, cur)) return false;
        return claim_now(stamp, d, tag);
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
    // a slot claimed by a delete of the same batch (kNone) has no state: the
    // waves may run while the deletes claim (its parent field never equals x)
    __device__ uint64_t peek(uint32_t y) const { return y != kNone ? state[y] : ~0ull; }
    __device__ bool edge(uint32_t x, uint32_t, uint32_t y, uint64_t cur) const {
        if (uint32_t(cur) != x) return false;
        // y's parent is x: y is a tree child reached once (twice only over
        // parallel edges), so its bit is almost always clear
        return bit_claim_now(aff, y);
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
    uint32_t n;  // touched rows >= n are incoming rows, which carry no rank
    PrScatterOp scatter;
    __device__ bool source(uint32_t v, uint32_t, Row &row, uint32_t &pay) const {
        if (v >= n) return false;
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
// Rows are read in quads: four consecutive slots with one 16-B load. Every
// row starts on a 4-slot boundary (its capacity, and so every allocation of
// the bump-allocated pools, is a power of two >= 4) and owns
// round_up4(deg) <= cap slots, so a quad never leaves its row's block. Slots
// of the last quad past `deg` are masked; their loaded words are replaced by
// the quad's first (live) slot before any use as an index.
struct WaveLds {
    uint32_t pref[kWave];  // exclusive scan of quad counts
    uint32_t v[kWave];
    uint32_t pay[kWave];
    uint32_t deg[kWave];
    uint32_t offlo[kWave];
    uint32_t offhi[kWave];
};

struct ExpandCfg {
    Pools pools;
    uint32_t split;     // rows with deg >= split go to tiles
    uint32_t tile_len;  // slots per tile (a multiple of 4)
    uint32_t stage;     // Stage whose counters this pass feeds
    StatShard *stats;
    // non-zero: a wave that takes fewer than 64 items raises the split in
    // proportion, up to this many slots (expand_range), so rows that its walk
    // can hold expand in this round instead of as tiles in the next one
    uint32_t adapt = 0;
    uint32_t pad = 0;
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

// ------------------------------------------------------------ appends
// Device-scope atomics on one address serialise where memory resolves them,
// so every append path reserves space for a whole wave (DirectOut) or for a
// whole workgroup (StagedOut): waves reserve in an LDS buffer with an LDS
// atomic and the workgroup moves the buffer to the global list with one
// global atomic when it finishes (out_flush). A reservation that does not
// fit in LDS spills its excess to the global list directly.

// A reservation of `total` consecutive elements: element i goes to LDS slot
// b+i when b+i < cap, else to global slot gb + (b+i) - max(b, cap).
struct Resv {
    uint32_t b, cap, gb;
};

template <class T>
__device__ __forceinline__ void resv_put(const Resv &r, uint32_t i, T *lds, T *glob, const T &val) {
    const uint32_t k = r.b + i;
    if (k < r.cap) lds[k] = val;
    else glob[r.gb + k - max(r.b, r.cap)] = val;
}

// Called by the whole wave with a wave-uniform total.
__device__ __forceinline__ Resv reserve_global(uint32_t *count, uint32_t total) {
    uint32_t g = 0;
    if (lane_id() == 0) g = atomicAdd(count, total);
    return Resv{0u, 0u, __shfl(g, 0, kWave)};
}

__device__ __forceinline__ Resv reserve_staged(uint32_t *lds_n, uint32_t cap, uint32_t *count,
                                               uint32_t total) {
    uint32_t b = 0, g = 0;
    if (lane_id() == 0) {
        b = atomicAdd(lds_n, total);
        const uint32_t lo = max(b, cap);
        if (b + total > lo) g = atomicAdd(count, b + total - lo);
    }
    return Resv{__shfl(b, 0, kWave), cap, __shfl(g, 0, kWave)};
}

struct DirectOut {
    Sink sink;
    TileSink ts;
};

template <uint32_t QCAP, uint32_t TCAP>
struct StageLds {
    uint32_t qn, qbase, tn, tbase;
    uint32_t q[QCAP];
    Tile t[TCAP];
    static_assert(QCAP % 2 == 0, "keeps the tile array 8-B aligned");
};

template <uint32_t QCAP, uint32_t TCAP>
struct StagedOut {
    Sink sink;
    TileSink ts;
    StageLds<QCAP, TCAP> *lds;
};

__device__ __forceinline__ Resv out_reserve_q(const DirectOut &o, uint32_t total) {
    return reserve_global(o.sink.count, total);
}
__device__ __forceinline__ Resv out_reserve_t(const DirectOut &o, uint32_t total) {
    return reserve_global(o.ts.count, total);
}
__device__ __forceinline__ uint32_t *out_qlds(const DirectOut &) { return nullptr; }
__device__ __forceinline__ Tile *out_tlds(const DirectOut &) { return nullptr; }
__device__ __forceinline__ void out_init(const DirectOut &) {}
__device__ __forceinline__ void out_flush(const DirectOut &) {}

template <uint32_t Q, uint32_t T>
__device__ __forceinline__ Resv out_reserve_q(const StagedOut<Q, T> &o, uint32_t total) {
    return reserve_staged(&o.lds->qn, Q, o.sink.count, total);
}
template <uint32_t Q, uint32_t T>
__device__ __forceinline__ Resv out_reserve_t(const StagedOut<Q, T> &o, uint32_t total) {
    return reserve_staged(&o.lds->tn, T, o.ts.count, total);
}
template <uint32_t Q, uint32_t T>
__device__ __forceinline__ uint32_t *out_qlds(const StagedOut<Q, T> &o) { return o.lds->q; }
template <uint32_t Q, uint32_t T>
__device__ __forceinline__ Tile *out_tlds(const StagedOut<Q, T> &o) { return o.lds->t; }

// Workgroup-level: every thread must call these (uniform control flow).
template <uint32_t Q, uint32_t T>
__device__ __forceinline__ void out_init(const StagedOut<Q, T> &o) {
    if (threadIdx.x == 0) o.lds->qn = o.lds->tn = 0u;
    __syncthreads();
}
template <uint32_t Q, uint32_t T>
__device__ void out_flush(const StagedOut<Q, T> &o) {
    StageLds<Q, T> *s = o.lds;
    __syncthreads();
    const uint32_t nq = min(s->qn, Q), nt = min(s->tn, T);
    if (threadIdx.x == 0) {
        s->qbase = nq ? atomicAdd(o.sink.count, nq) : 0u;
        s->tbase = nt ? atomicAdd(o.ts.count, nt) : 0u;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nq; i += blockDim.x) o.sink.list[s->qbase + i] = s->q[i];
    for (uint32_t i = threadIdx.x; i < nt; i += blockDim.x) o.ts.tiles[s->tbase + i] = s->t[i];
}

// Appends v[k] for every set f[k] of every lane (wave-uniform call).
template <int K, class Out>
__device__ __forceinline__ void out_push(const Out &o, const bool (&f)[K], const uint32_t (&v)[K]) {
    uint64_t m[K];
    uint32_t total = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        m[k] = __ballot(f[k]);
        total += uint32_t(__popcll(m[k]));
    }
    if (total == 0) return;
    const Resv r = out_reserve_q(o, total);
    uint32_t *lq = out_qlds(o);
    const uint64_t lt = __lanemask_lt();
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (f[k]) resv_put(r, run + uint32_t(__popcll(m[k] & lt)), lq, o.sink.list, v[k]);
        run += uint32_t(__popcll(m[k]));
    }
}

// ------------------------------------------------- workgroup reservations
template <class T>
__device__ __forceinline__ T wave_incl_scan_t(T x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T y = __shfl_up(x, d, kWave);
        if (lane >= uint32_t(d)) x += y;
    }
    return x;
}

constexpr uint32_t kMaxWavesPerBlock = 16;

// Exclusive prefix of x over the workgroup; the workgroup's total is
// reserved from *counter with one atomic. Every thread must call it;
// `scratch` is kMaxWavesPerBlock+1 LDS words of type T.
template <class T>
__device__ T block_reserve(T x, T *counter, T *scratch) {
    const uint32_t lane = lane_id(), w = wave_in_block(), nw = blockDim.x / kWave;
    const T incl = wave_incl_scan_t<T>(x);
    if (lane == kWave - 1) scratch[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (uint32_t i = 0; i < nw; ++i) {
            const T t = scratch[i];
            scratch[i] = run;
            run += t;
        }
        scratch[kMaxWavesPerBlock] = run ? atomicAdd(counter, run) : T(0);
    }
    __syncthreads();
    const T r = scratch[kMaxWavesPerBlock] + scratch[w] + incl - x;
    __syncthreads();  // scratch is reused by the next call
    return r;
}

// block_reserve for K counters at once: one scan pass, then K threads of
// wave 0 issue the K atomics together (one memory round trip per call).
// `scratch` is K*(kMaxWavesPerBlock+1) LDS words. Every thread must call it.
template <int K>
__device__ void block_reserve_n(const uint32_t (&x)[K], uint32_t *const (&counter)[K], uint32_t (&out)[K],
                                uint32_t *scratch) {
    constexpr uint32_t S = kMaxWavesPerBlock + 1;
    const uint32_t lane = lane_id(), w = wave_in_block(), nw = blockDim.x / kWave;
    uint32_t incl[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        incl[k] = wave_incl_scan(x[k]);
        if (lane == kWave - 1) scratch[k * S + w] = incl[k];
    }
    __syncthreads();
    if (threadIdx.x < uint32_t(K)) {
        const uint32_t k = threadIdx.x;
        uint32_t run = 0;
        for (uint32_t i = 0; i < nw; ++i) {
            const uint32_t t = scratch[k * S + i];
            scratch[k * S + i] = run;
            run += t;
        }
        scratch[k * S + kMaxWavesPerBlock] = run ? atomicAdd(counter[k], run) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = scratch[k * S + kMaxWavesPerBlock] + scratch[k * S + w] + incl[k] - x[k];
    __syncthreads();
}

// One-element-per-thread append through block_reserve.
__device__ __forceinline__ void block_append(bool f, uint32_t v, uint32_t *list, uint32_t *count,
                                             uint32_t *scratch) {
    const uint32_t at = block_reserve<uint32_t>(f ? 1u : 0u, count, scratch);
    if (f) list[at] = v;
}

// ------------------------------------------------------------ statistics
__device__ __forceinline__ StatShard *stat_shard(StatShard *s) {
    const uint32_t w = blockIdx.x * (blockDim.x / kWave) + wave_in_block();
    return s + (w & (kStatShards - 1));
}

// Adds one wave-total to a shard field (lane 0; the value must be wave-uniform).
__device__ __forceinline__ void stat_add(uint64_t *field, uint64_t v) {
    if (v && lane_id() == 0) atomicAdd((unsigned long long *)field, (unsigned long long)v);
}

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

// Loads quad `qi` of a row of `deg` slots: the live slot words, with dead
// slots aliased to slot 0 of the quad; returns the number of live slots.
__device__ __forceinline__ uint32_t load_quad(const uint32_t *row, uint32_t qi, uint32_t deg,
                                              uint32_t (&d)[4]) {
    const uint4 x = *reinterpret_cast<const uint4 *>(row + 4ull * qi);
    const uint32_t live = min(deg - 4u * qi, 4u);
    d[0] = x.x;
    d[1] = live > 1 ? x.y : x.x;
    d[2] = live > 2 ? x.z : x.x;
    d[3] = live > 3 ? x.w : x.x;
    return live;
}

// load_quad for rows streamed once per pass (PageRank's incoming-row
// sweeps): a nontemporal load, so the stream does not push the gathered
// contributions out of L2.
typedef uint32_t u32x4_nt __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint32_t load_quad_stream(const uint32_t *row, uint32_t qi, uint32_t deg,
                                                     uint32_t (&d)[4]) {
    const u32x4_nt x = __builtin_nontemporal_load(reinterpret_cast<const u32x4_nt *>(row + 4ull * qi));
    const uint32_t live = min(deg - 4u * qi, 4u);
    d[0] = x.x;
    d[1] = live > 1 ? x.y : x.x;
    d[2] = live > 2 ? x.z : x.x;
    d[3] = live > 3 ? x.w : x.x;
    return live;
}

// Expands the 64 items [base, base+64) of `items` (or the identity range when
// items == nullptr) that are below `count`. Each lane takes QPL quads of the
// wave's concatenated rows per step, so a step covers 256*QPL slots.
template <int QPL, class Op, class Out>
__device__ void expand_chunk(const Op &op, const ExpandCfg &cfg, const uint32_t *items,
                             uint32_t base, uint32_t count, WaveLds &L, const Out &out,
                             Counters &ctr) {
    const uint32_t lane = lane_id();
    const uint32_t i = base + lane;
    uint32_t v = 0, pay = 0, deg = 0;
    uint64_t off = 0;
    if (i < count) {
        v = items ? items[i] : i;
        Row row;
        if (v != kNone && op.source(v, i, row, pay)) {
            deg = row.deg;
            off = row.off;
        }
    }
    ctr.scanned += deg;
    ctr.vexp += deg ? 1 : 0;
    if (off & kHostTier) ctr.host += deg;
    // long rows become tiles; one reservation per wave
    const uint32_t nt = deg >= cfg.split ? (deg + cfg.tile_len - 1) / cfg.tile_len : 0u;
    const uint32_t tincl = wave_incl_scan(nt);
    const uint32_t ttotal = __shfl(tincl, kWave - 1, kWave);
    if (ttotal) {
        const Resv r = out_reserve_t(out, ttotal);
        Tile *lt = out_tlds(out);
        for (uint32_t j = 0; j < nt; ++j) {
            Tile t;
            t.start = off + uint64_t(j) * cfg.tile_len;
            t.v = v;
            t.len = min(cfg.tile_len, deg - j * cfg.tile_len);
            t.payload = pay;
            t.pad = 0;
            resv_put(r, tincl - nt + j, lt, out.ts.tiles, t);
        }
        if (nt) deg = 0;
    }
    const uint32_t nq = (deg + 3u) >> 2;
    const uint32_t incl = wave_incl_scan(nq);
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    if (total == 0) return;
    L.pref[lane] = incl - nq;
    L.v[lane] = v;
    L.pay[lane] = pay;
    L.deg[lane] = deg;
    L.offlo[lane] = uint32_t(off);
    L.offhi[lane] = uint32_t(off >> 32);
    wave_lds_sync();
    for (uint32_t k0 = 0; k0 < total; k0 += kWave * QPL) {
        uint32_t d[QPL][4], own[QPL], live[QPL];
        // owner searches and loads are unconditional (quad indices clamped to
        // the last quad) so that all QPL loads are in flight together
#pragma unroll
        for (int q = 0; q < QPL; ++q) {
            const uint32_t k = k0 + uint32_t(q) * kWave + lane;
            const uint32_t kk = k < total ? k : total - 1;
            own[q] = owner_of(L.pref, kk);
            const uint64_t o = (uint64_t(L.offhi[own[q]]) << 32) | L.offlo[own[q]];
            live[q] = load_quad(cfg.pools.row(o), kk - L.pref[own[q]], L.deg[own[q]], d[q]);
            if (k >= total) live[q] = 0;
        }
        uint64_t cur[QPL][4];
#pragma unroll
        for (int q = 0; q < QPL; ++q)
#pragma unroll
            for (int s = 0; s < 4; ++s) cur[q][s] = op.peek(d[q][s]);
        bool push[QPL * 4];
        uint32_t dst[QPL * 4];
#pragma unroll
        for (int q = 0; q < QPL; ++q) {
            const uint32_t u = L.v[own[q]], p = L.pay[own[q]];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                push[q * 4 + s] = uint32_t(s) < live[q] && op.edge(u, p, d[q][s], cur[q][s]);
                dst[q * 4 + s] = d[q][s];
            }
        }
        out_push<QPL * 4>(out, push, dst);
    }
    wave_lds_sync();  // LDS is reused by the next chunk of this wave
}

template <int QPL, class Op, class Out>
__device__ void expand_tile(const Op &op, const ExpandCfg &cfg, const Tile &t, const Out &out) {
    uint32_t pay;
    if (t.len == 0 || !op.tile_source(t, pay)) return;
    const uint32_t *p = cfg.pools.row(t.start);  // t.start is a multiple of 4
    const uint32_t lane = lane_id();
    const uint32_t nq = (t.len + 3u) >> 2;
    for (uint32_t k0 = 0; k0 < nq; k0 += kWave * QPL) {
        uint32_t d[QPL][4], live[QPL];
#pragma unroll
        for (int q = 0; q < QPL; ++q) {
            const uint32_t k = k0 + uint32_t(q) * kWave + lane;
            live[q] = load_quad(p, k < nq ? k : nq - 1, t.len, d[q]);
            if (k >= nq) live[q] = 0;
        }
        uint64_t cur[QPL][4];
#pragma unroll
        for (int q = 0; q < QPL; ++q)
#pragma unroll
            for (int s = 0; s < 4; ++s) cur[q][s] = op.peek(d[q][s]);
        bool push[QPL * 4];
        uint32_t dst[QPL * 4];
#pragma unroll
        for (int q = 0; q < QPL; ++q)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                push[q * 4 + s] = uint32_t(s) < live[q] && op.edge(t.v, pay, d[q][s], cur[q][s]);
                dst[q * 4 + s] = d[q][s];
            }
        out_push<QPL * 4>(out, push, dst);
    }
}

// Grid-strided driver over a list (or identity range) and then a tile list;
// any block size that is a multiple of 64.
// A wave takes `per` consecutive items at a time: 64 when the list is long,
// fewer (a power of two) when the list is shorter than 64 items per wave of
// the grid, so that small frontiers still spread over every wave.
template <int QPL, class Op, class Out>
__device__ void expand_range(const Op &op, const ExpandCfg &cfg, const uint32_t *items,
                             uint32_t count, const Out &out, WaveLds &L, Counters &ctr) {
    const uint32_t wpb = blockDim.x / kWave;
    const uint32_t gwave = blockIdx.x * wpb + wave_in_block();
    const uint32_t nwaves = gridDim.x * wpb;
    uint32_t per = kWave;
    while (per > 1 && uint64_t(per / 2) * nwaves >= count) per >>= 1;
    // A wave walks at most per * split slots of short rows; with fewer items
    // per wave the same bound admits longer rows. A row cut into tiles waits
    // for the next round, so in a chain of frontier rounds every row kept
    // whole here saves its children a round.
    ExpandCfg c = cfg;
    if (cfg.adapt) c.split = max(cfg.split, min(cfg.adapt, cfg.split * (kWave / per)));
    for (uint64_t base = uint64_t(gwave) * per; base < count; base += uint64_t(nwaves) * per)
        expand_chunk<QPL>(op, c, items, uint32_t(base), uint32_t(min(uint64_t(count), base + per)), L,
                          out, ctr);
}

template <int QPL, class Op, class Out>
__device__ void expand_tiles(const Op &op, const ExpandCfg &cfg, const Tile *tiles,
                             uint32_t ntiles, const Out &out) {
    const uint32_t wpb = blockDim.x / kWave;
    const uint32_t gwave = blockIdx.x * wpb + wave_in_block();
    const uint32_t nwaves = gridDim.x * wpb;
    for (uint32_t t = gwave; t < ntiles; t += nwaves) {
        const Tile tile = tiles[t];
        expand_tile<QPL>(op, cfg, tile, out);
    }
}

// Adds this wave's counters to its statistics shard (one atomic per wave and
// non-zero counter).
__device__ __forceinline__ void flush_counters(Counters &ctr, StatShard *stats, uint32_t stage) {
    const uint64_t s = wave_sum64(ctr.scanned), h = wave_sum64(ctr.host), x = wave_sum64(ctr.vexp);
    StatShard *sh = stat_shard(stats);
    stat_add(&sh->edges[stage], s);
    stat_add(&sh->host[stage], h);
    stat_add(&sh->vexp[stage], x);
    ctr = Counters{};
}

// Round progress for the host driver: every round launch posts exactly one
// word to pinned host memory, (tag << 1) | has_work, where tag = chain epoch
// and launch sequence number; workgroup 0 posts it when the round starts.
// The host reads nothing else through it (results are read after a stream
// synchronisation), so the store needs no ordering, only system scope.
__device__ __forceinline__ void post_progress(uint64_t *prog, uint64_t tag, bool more) {
    __hip_atomic_store(prog, (tag << 1) | (more ? 1ull : 0ull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace dev
}  // namespace cig
