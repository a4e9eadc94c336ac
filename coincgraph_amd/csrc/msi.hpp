// Incremental BFS / SSSP from up to 64 sources at once.
//
// The reference maintains one source per engine (InitGraph takes the source,
// framework.cuh:991-1041; the batch path is update_tree_del / update_tree_add,
// framework.cuh:2139-2348, 2752-2894). Here one engine keeps k <= 64 sources
// current through every batch with the same steps as the single-source path
// (mutation, invalidation of parent-tree dependents, pull repair, insert
// seeding, push relaxation), each step done for all sources in one pass over
// the rows it touches:
//   st[v*k + s]  8 B  value << 32 | parent of v for source s (vertex-major:
//                     the k states of a vertex share cache lines)
//   act[v]       8 B  sources whose value at v improved, pending expansion
//   snap[v]      8 B  the round's snapshot of act (or of pinv)
//   affm[v]      8 B  sources for which v is affected in the current call
//   pinv[v]      8 B  affected bits not yet expanded by an invalidation wave
// Every per-edge step loops over the set bits of a 64-bit source mask; the
// row itself is read once for all of them.
#pragma once

namespace cig {
namespace dev {

struct MsiArgs {
    Ctl *ctl;
    const Row *rows;
    const Row *in_rows;  // incoming rows (rows + n, or rows for a symmetric graph)
    unsigned long long *st;
    unsigned long long *act, *snap, *affm, *pinv;
    uint32_t *stamp;
    uint32_t *q0, *q1;
    Tile *t0, *t1;
    uint32_t *alist;
    ExpandCfg cfg;
    const cig_edge *dels;
    uint32_t nd;
    const cig_edge *adds;
    uint32_t na;
    const uint8_t *add_skip, *del_skip;  // grouped batches: cancelled records (may be null)
    uint32_t n, k;
    uint64_t seed;
};

constexpr unsigned long long kMsiNone = ~0ull;  // value kInf, parent kNone

__device__ __forceinline__ uint32_t msi_bit(unsigned long long &m) {
    const uint32_t s = uint32_t(__ffsll(m) - 1);
    m &= m - 1;
    return s;
}

// Sources: state (0, none) and the first frontier. The host has set every
// state word to kMsiNone and every mask to 0.
__global__ void k_msi_init(MsiArgs a, SrcList sources) {
    const uint32_t r = a.ctl->round;
    for (uint32_t s = gthread(); s < a.k; s += gthreads()) {
        const uint32_t v = sources.v[s];
        a.st[uint64_t(v) * a.k + s] = uint64_t(kNone);  // value 0, no parent
        atomicOr(&a.act[v], 1ull << s);
        if (claim(a.stamp, v, r)) {
            const uint32_t at = atomicAdd(&a.ctl->qcount[r % 3], 1u);
            ((r & 1) ? a.q1 : a.q0)[at] = v;
        }
    }
}

// Marks bits `m` of y affected; returns the newly affected ones, which are
// also left pending for the next invalidation wave.
__device__ __forceinline__ unsigned long long msi_affect(unsigned long long *affm, unsigned long long *pinv,
                                                         uint32_t y, unsigned long long m) {
    const unsigned long long gained = m & ~atomicOr(&affm[y], m);
    if (gained) atomicOr(&pinv[y], gained);
    return gained;
}

// reset_del_edges (driver.cuh:371-411) per source: a deleted (u,v) whose v
// has parent u for source s makes (v, s) affected. Parallel copies of (u,v)
// are treated conservatively, as in the single-source path.
__global__ void k_msi_inval_seed(MsiArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    const uint32_t r = a.ctl->round;
    uint32_t *q = (r & 1) ? a.q1 : a.q0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.nd; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        bool push = false;
        uint32_t v = 0;
        if (i < a.nd) {
            const cig_edge e = a.dels[i];
            if (e.u < a.n && e.v < a.n && !(a.del_skip && a.del_skip[i])) {
                const unsigned long long *sv = a.st + uint64_t(e.v) * a.k;
                unsigned long long m = 0;
                for (uint32_t s = 0; s < a.k; ++s)
                    if (uint32_t(sv[s]) == e.u) m |= 1ull << s;
                if (m && msi_affect(a.affm, a.pinv, e.v, m)) {
                    v = e.v;
                    push = claim(a.stamp, v, r);
                }
            }
        }
        block_append(push, v, q, &a.ctl->qcount[r % 3], scratch);
    }
}

// Invalidation wave (RelaxCTADB_del + PushFunctorDB_Del, sync_push_dd.cuh:
// 1220-1303): for each source s newly affected at x, a child y whose parent
// for s is x becomes affected for s.
struct MsiInvalOp {
    const Row *rows;
    const unsigned long long *snap;
    const unsigned long long *st;
    unsigned long long *affm, *pinv;
    uint32_t *stamp;
    uint32_t k, tag;
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
    __device__ bool edge(uint32_t x, uint32_t, uint32_t y, uint64_t) const {
        // a slot claimed by a delete of the same batch (the claims may run
        // beside the waves) holds kNone and no vertex
        if (y == kNone) return false;
        const unsigned long long *sy = st + uint64_t(y) * k;
        unsigned long long m = 0;
        for (unsigned long long b = snap[x]; b;) {
            const uint32_t s = msi_bit(b);
            if (uint32_t(sy[s]) == x) m |= 1ull << s;
        }
        if (!m || !msi_affect(affm, pinv, y, m)) return false;
        return claim(stamp, y, tag);
    }
};

// Push relaxation for every source pending at the expanding vertex
// (SyncPushDDB relax + AccumulateBuffer per source). Eight sources' loads are
// issued together; an improvement is a 64-bit minimum of value << 32 |
// parent, so equal values keep the parent already there.
template <int W>
struct MsiRelaxOp {
    const Row *rows;
    unsigned long long *st;
    unsigned long long *act;
    const unsigned long long *snap;
    uint32_t *stamp;
    uint32_t k, tag;
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
    __device__ bool edge(uint32_t u, uint32_t, uint32_t w, uint64_t) const {
        const uint32_t wt = weight<W>(u, w, seed);
        const unsigned long long *su = st + uint64_t(u) * k;
        unsigned long long *sw = st + uint64_t(w) * k;
        unsigned long long gained = 0;
        for (unsigned long long m = snap[u]; m;) {
            constexpr int kB = 8;
            uint32_t s[kB];
            unsigned long long du[kB], dw[kB];
            int cnt = 0;
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                s[j] = m ? uint32_t(__ffsll(m) - 1) : 0u;
                cnt += m ? 1 : 0;
                m &= m - 1;
            }
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                du[j] = j < cnt ? su[s[j]] : kMsiNone;
                dw[j] = j < cnt ? sw[s[j]] : 0ull;
            }
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const uint64_t c = (du[j] >> 32) + wt;
                if (j < cnt && c < (dw[j] >> 32)) {  // also rejects c >= kInf
                    const unsigned long long nv = (c << 32) | u;
                    const unsigned long long old = atomicMin(&sw[s[j]], nv);
                    if ((old >> 32) > c) gained |= 1ull << s[j];
                }
            }
        }
        if (!gained) return false;
        atomicOr(&act[w], gained);
        return claim(stamp, w, tag);
    }
};

// Round round+seq-1, phase 1 (relaxation or invalidation): snapshot and
// clear the pending masks `from` of the frontier; workgroup 0 posts progress.
__global__ void k_msi_consume(MsiArgs a, unsigned long long *from, uint64_t *prog, uint64_t tag, uint32_t seq) {
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
        a.snap[v] = atomicExch(&from[v], 0ull);
    }
}

template <int W>
__device__ __forceinline__ MsiRelaxOp<W> msi_relax_op(const MsiArgs &a, uint32_t tag) {
    return MsiRelaxOp<W>{a.rows, a.st, a.act, a.snap, a.stamp, a.k, tag, a.seed};
}

// Phase 2: expand the frontier rows (long rows become tiles for phase 3).
// INV selects the invalidation wave, else the relaxation of weight kind W.
template <int W, bool INV, int QPL>
__global__ __launch_bounds__(kBlock) void k_msi_expand(MsiArgs a, uint32_t seq) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ RoundStage stage;
    Ctl *c = a.ctl;
    const uint32_t r = c->round + seq - 1;
    const uint32_t nq = c->qcount[r % 3];
    if (nq == 0) return;
    Counters ctr;
    const uint32_t *qin = (r & 1) ? a.q1 : a.q0;
    const StagedOut<2048, 128> out{Sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]},
                                   TileSink{a.t0, &c->scount}, &stage};
    out_init(out);
    if constexpr (INV) {
        const MsiInvalOp op{a.rows, a.snap, a.st, a.affm, a.pinv, a.stamp, a.k, r + 1};
        expand_range<QPL>(op, a.cfg, qin, nq, out, lds[wave_in_block()], ctr);
    } else {
        expand_range<QPL>(msi_relax_op<W>(a, r + 1), a.cfg, qin, nq, out, lds[wave_in_block()], ctr);
    }
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
}

// Phase 3: the tiles cut in phase 2, with the same snapshot.
template <int W, bool INV, int QPL>
__global__ __launch_bounds__(kBlock) void k_msi_tiles(MsiArgs a, uint32_t seq) {
    __shared__ RoundStage stage;
    Ctl *c = a.ctl;
    const uint32_t r = c->round + seq - 1;
    const uint32_t nt = c->scount;
    if (nt == 0) return;
    const StagedOut<2048, 128> out{Sink{(r & 1) ? a.q0 : a.q1, &c->qcount[(r + 1) % 3]},
                                   TileSink{nullptr, &c->scount}, &stage};
    out_init(out);
    if constexpr (INV) {
        const MsiInvalOp op{a.rows, a.snap, a.st, a.affm, a.pinv, a.stamp, a.k, r + 1};
        expand_tiles<QPL>(op, a.cfg, a.t0, nt, out);
    } else {
        expand_tiles<QPL>(msi_relax_op<W>(a, r + 1), a.cfg, a.t0, nt, out);
    }
    out_flush(out);
}

// reset_affected_values (driver.cuh:418-433) per source, over every vertex:
// each affected (v, s) restarts from (kInf, none), and v joins the affected
// list that the repair and the end-of-call clear walk.
__global__ void k_msi_reset(MsiArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.n; base += gthreads()) {
        const uint32_t v = base + threadIdx.x;
        unsigned long long m = v < a.n ? a.affm[v] : 0ull;
        const bool hit = m != 0ull;
        unsigned long long *sv = a.st + uint64_t(v) * a.k;
        while (m) sv[msi_bit(m)] = kMsiNone;
        block_append(hit, v, a.alist, &a.ctl->acount, scratch);
    }
}

// GpuAffectedPullRelax (framework.cuh:137-204) per source: each affected
// (y, s) takes the minimum over in-neighbours x not affected for s.
template <int W>
struct MsiPullOp {
    const Row *in_rows;
    unsigned long long *st;
    const unsigned long long *affm;
    uint32_t k;
    uint64_t seed;
    __device__ bool source(uint32_t y, uint32_t, Row &row, uint32_t &pay) const {
        row = in_rows[y];
        pay = 0;
        return row.deg != 0;
    }
    __device__ bool tile_source(const Tile &, uint32_t &pay) const {
        pay = 0;
        return true;
    }
    __device__ uint64_t peek(uint32_t x) const { return affm[x]; }
    __device__ bool edge(uint32_t y, uint32_t, uint32_t x, uint64_t ax) const {
        const uint32_t wt = weight<W>(x, y, seed);
        const unsigned long long *sx = st + uint64_t(x) * k;
        unsigned long long *sy = st + uint64_t(y) * k;
        for (unsigned long long m = affm[y] & ~ax; m;) {
            const uint32_t s = msi_bit(m);
            const uint64_t c = (sx[s] >> 32) + wt;
            if (c < (sy[s] >> 32)) atomicMin(&sy[s], (c << 32) | x);  // also rejects c >= kInf
        }
        return false;
    }
};

template <int W, int QPL>
__global__ __launch_bounds__(kBlock) void k_msi_repair(MsiArgs a) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ StageLds<2, 128> stage;
    Ctl *c = a.ctl;
    const uint32_t na = ld_acquire(&c->acount);
    if (na == 0) return;
    const MsiPullOp<W> op{a.in_rows, a.st, a.affm, a.k, a.seed};
    Counters ctr;
    const StagedOut<2, 128> out{Sink{nullptr, &c->scount}, TileSink{a.t0, &c->scount}, &stage};
    out_init(out);
    expand_range<QPL>(op, a.cfg, a.alist, na, out, lds[wave_in_block()], ctr);
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
}

template <int W, int QPL>
__global__ __launch_bounds__(kBlock) void k_msi_repair_tiles(MsiArgs a) {
    Ctl *c = a.ctl;
    const uint32_t nt = ld_acquire(&c->scount);
    if (nt == 0) return;
    const MsiPullOp<W> op{a.in_rows, a.st, a.affm, a.k, a.seed};
    expand_tiles<QPL>(op, a.cfg, a.t0, nt, DirectOut{Sink{nullptr, &c->scount}, TileSink{nullptr, &c->scount}});
}

// After the repair: every affected (y, s) that has a value again is pending
// expansion, and y joins the frontier of the current round.
__global__ void k_msi_repair_done(MsiArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    const uint32_t cnt = ld_acquire(&a.ctl->acount);
    const uint32_t r = a.ctl->round;
    uint32_t *q = (r & 1) ? a.q1 : a.q0;
    if (threadIdx.x == 0 && blockIdx.x == 0) a.ctl->scount = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < cnt; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        bool push = false;
        uint32_t y = 0;
        if (i < cnt) {
            y = a.alist[i];
            const unsigned long long *sy = a.st + uint64_t(y) * a.k;
            unsigned long long f = 0;
            for (unsigned long long m = a.affm[y]; m;) {
                const uint32_t s = msi_bit(m);
                if ((sy[s] >> 32) != kInf) f |= 1ull << s;
            }
            if (f) {
                atomicOr(&a.act[y], f);
                push = claim(a.stamp, y, r);
            }
        }
        block_append(push, y, q, &a.ctl->qcount[r % 3], scratch);
    }
}

// SeedAddedEdgeFrontier (framework.cuh:238-294) per source: an inserted
// (u,v) offers value_s(u) + w to v for every source s not affected at u (an
// affected u re-expands its whole row once repaired).
template <int W>
__global__ void k_msi_insert_seed(MsiArgs a, uint32_t skip_affected) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    const uint32_t r = a.ctl->round;
    uint32_t *q = (r & 1) ? a.q1 : a.q0;
    const unsigned long long all = a.k == 64 ? ~0ull : (1ull << a.k) - 1;
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.na; base += gthreads()) {
        const uint32_t i = base + threadIdx.x;
        bool push = false;
        uint32_t v = 0;
        if (i < a.na) {
            const cig_edge e = a.adds[i];
            if (e.u < a.n && e.v < a.n && !(a.add_skip && a.add_skip[i])) {
                const uint32_t wt = weight<W>(e.u, e.v, a.seed);
                const unsigned long long *su = a.st + uint64_t(e.u) * a.k;
                unsigned long long *sv = a.st + uint64_t(e.v) * a.k;
                unsigned long long gained = 0;
                for (unsigned long long m = skip_affected ? all & ~a.affm[e.u] : all; m;) {
                    const uint32_t s = msi_bit(m);
                    const uint64_t c = (su[s] >> 32) + wt;
                    if (c < (sv[s] >> 32)) {
                        const unsigned long long old = atomicMin(&sv[s], (c << 32) | e.u);
                        if ((old >> 32) > c) gained |= 1ull << s;
                    }
                }
                if (gained) {
                    v = e.v;
                    atomicOr(&a.act[v], gained);
                    push = claim(a.stamp, v, r);
                }
            }
        }
        block_append(push, v, q, &a.ctl->qcount[r % 3], scratch);
    }
}

// End of a call: the affected marks of the listed vertices go back to 0.
__global__ void k_msi_clear(MsiArgs a) {
    const uint32_t cnt = ld_acquire(&a.ctl->acount);
    for (uint32_t i = gthread(); i < cnt; i += gthreads()) {
        const uint32_t v = a.alist[i];
        a.affm[v] = 0ull;
        a.pinv[v] = 0ull;
    }
}

// Source-major result: out[s * n + v] = value (field 0) or parent (field 1).
__global__ void k_msi_gather(const unsigned long long *st, uint32_t n, uint32_t k, uint32_t field, uint32_t *out) {
    const uint64_t total = uint64_t(n) * k;
    for (uint64_t i = gthread(); i < total; i += gthreads()) {
        const uint32_t s = uint32_t(i / n), v = uint32_t(i % n);
        const unsigned long long x = st[uint64_t(v) * k + s];
        out[i] = field == 0 ? uint32_t(x >> 32) : uint32_t(x);
    }
}

}  // namespace dev
}  // namespace cig
