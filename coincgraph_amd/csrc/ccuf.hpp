// Min-label CC on an undirected graph by union-find (the reference's union
// repair, cc_union_repair.cuh:9-137, used instead of label propagation with
// CIG_CFG_CC_UNION_REPAIR).
//
// Invariant between calls: uf[v] = label[v], the smallest vertex of v's
// component, and uf[label] = label. A delete can split only the component it
// lies in, and no edge leaves a component, so every component a delete
// touches is re-derived on its own: its vertices restart as singletons and
// unite along their rows. Roots are linked toward the smaller id, so a root
// is always the smallest vertex of its set. An insert unites two components.
// The state word keeps the label in its high half; no parent is kept.
#pragma once

namespace cig {
namespace dev {

struct UfArgs {
    Ctl *ctl;
    const Row *rows;
    uint64_t *state;
    uint32_t *uf;
    uint32_t *roots;  // snapshot of the sampled roots of the affected vertices
    uint32_t *marks;  // bitmap: labels of the components the call's deletes touch
    uint32_t *alist;  // the affected vertices
    Tile *tiles;
    ExpandCfg cfg;
    const cig_edge *dels;
    uint32_t nd;
    const cig_edge *adds;
    uint32_t na;
    const uint8_t *add_skip, *del_skip;
    uint32_t n;
};

__device__ __forceinline__ uint32_t uf_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of v, halving the path on the way (Root, cc_union_repair.cuh:9-22).
__device__ __forceinline__ uint32_t uf_find(uint32_t *uf, uint32_t v) {
    uint32_t p = uf_load(uf + v);
    while (p != v) {
        const uint32_t gp = uf_load(uf + p);
        if (gp != p) atomicCAS(uf + v, p, gp);
        v = p;
        p = uf_load(uf + v);
    }
    return v;
}

// Links the larger root under the smaller one; a stale root retries.
__device__ __forceinline__ void uf_unite(uint32_t *uf, uint32_t a, uint32_t b) {
    for (;;) {
        a = uf_find(uf, a);
        b = uf_find(uf, b);
        if (a == b) return;
        const uint32_t hi = max(a, b), lo = min(a, b);
        if (atomicCAS(uf + hi, hi, lo) == hi) return;
    }
}

// Initial state: every vertex alone and affected (the whole graph is
// derived like a deleted-into component).
__global__ void k_uf_init(UfArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.n; base += gthreads()) {
        const uint32_t v = base + threadIdx.x;
        if (v < a.n) a.uf[v] = v;
        block_append(v < a.n, v, a.alist, &a.ctl->acount, scratch);
    }
}

// The labels of the components the deletes touch (both endpoints of an
// undirected edge carry the same label).
__global__ void k_uf_mark(UfArgs a) {
    for (uint32_t i = gthread(); i < a.nd; i += gthreads()) {
        const cig_edge e = a.dels[i];
        if (e.u >= a.n || e.v >= a.n || (a.del_skip && a.del_skip[i])) continue;
        const uint32_t l = uint32_t(a.state[e.u] >> 32);
        if (!bit_test(a.marks, l)) atomicOr(&a.marks[l >> 5], 1u << (l & 31));
    }
}

// Every vertex of a marked component restarts alone and is listed.
__global__ void k_uf_collect(UfArgs a) {
    __shared__ uint32_t scratch[kMaxWavesPerBlock + 1];
    for (uint32_t base = blockIdx.x * blockDim.x; base < a.n; base += gthreads()) {
        const uint32_t v = base + threadIdx.x;
        bool hit = false;
        if (v < a.n) {
            hit = bit_test(a.marks, uint32_t(a.state[v] >> 32));
            if (hit) a.uf[v] = v;
        }
        block_append(hit, v, a.alist, &a.ctl->acount, scratch);
    }
}

// Sample (cc_union_repair.cuh:41-60): each affected vertex unites with its
// first two neighbours, which joins most of a giant component cheaply.
__global__ void k_uf_sample(UfArgs a) {
    const uint32_t cnt = ld_acquire(&a.ctl->acount);
    for (uint32_t i = gthread(); i < cnt; i += gthreads()) {
        const uint32_t u = a.alist[i];
        const Row row = a.rows[u];
        const uint32_t *p = a.cfg.pools.row(row.off);
        for (uint32_t j = 0; j < 2 && j < row.deg; ++j) uf_unite(a.uf, u, p[j]);
    }
}

// Snapshot (cc_union_repair.cuh:62-69): the sampled root of every affected vertex.
__global__ void k_uf_snapshot(UfArgs a) {
    const uint32_t cnt = ld_acquire(&a.ctl->acount);
    for (uint32_t i = gthread(); i < cnt; i += gthreads()) {
        const uint32_t u = a.alist[i];
        a.roots[u] = uf_find(a.uf, u);
    }
}

// SelectComponent (cc_union_repair.cuh:71-93): the most frequent root among
// 256 evenly spaced affected vertices; its members skip hooking.
__global__ void k_uf_select(UfArgs a) {
    __shared__ uint32_t cand[256], votes[256];
    const uint32_t t = threadIdx.x;
    const uint32_t cnt = ld_acquire(&a.ctl->acount);
    cand[t] = cnt ? a.roots[a.alist[uint32_t((uint64_t(t) * cnt) / 256)]] : kNone;
    __syncthreads();
    uint32_t k = 0;
    for (uint32_t j = 0; j < 256; ++j) k += cand[j] == cand[t];
    votes[t] = k;
    __syncthreads();
    if (t == 0) {
        uint32_t best = 0;
        for (uint32_t j = 1; j < 256; ++j)
            if (votes[j] > votes[best]) best = j;
        a.ctl->ufsel[0] = cand[best];
        a.ctl->ufsel[1] = votes[best];
    }
}

// Hook (cc_union_repair.cuh:95-124) as an expansion pass: every edge of an
// affected vertex outside the selected component unites its two ends when
// their sampled roots differ. Each undirected edge is taken from its smaller
// end, except a cut edge into the selected component, whose selected end
// does not expand.
struct UfHookOp {
    const Row *rows;
    uint32_t *uf;
    const uint32_t *roots;
    uint32_t sel;
    __device__ bool source(uint32_t u, uint32_t, Row &row, uint32_t &pay) const {
        if (roots[u] == sel) return false;
        row = rows[u];
        pay = 0;
        return row.deg != 0;
    }
    __device__ bool tile_source(const Tile &, uint32_t &pay) const {
        pay = 0;
        return true;
    }
    __device__ uint64_t peek(uint32_t v) const { return roots[v]; }
    __device__ bool edge(uint32_t u, uint32_t, uint32_t v, uint64_t rv) const {
        const uint32_t ru = roots[u];
        if (ru != uint32_t(rv) && (u < v || uint32_t(rv) == sel)) uf_unite(uf, u, v);
        return false;
    }
};

template <int QPL>
__global__ __launch_bounds__(kBlock) void k_uf_hook(UfArgs a) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ StageLds<2, 128> stage;
    Ctl *c = a.ctl;
    const uint32_t cnt = ld_acquire(&c->acount);
    if (cnt == 0) return;
    const UfHookOp op{a.rows, a.uf, a.roots, c->ufsel[0]};
    Counters ctr;
    const StagedOut<2, 128> out{Sink{nullptr, &c->scount}, TileSink{a.tiles, &c->scount}, &stage};
    out_init(out);
    expand_range<QPL>(op, a.cfg, a.alist, cnt, out, lds[wave_in_block()], ctr);
    flush_counters(ctr, a.cfg.stats, a.cfg.stage);
    out_flush(out);
}

template <int QPL>
__global__ __launch_bounds__(kBlock) void k_uf_hook_tiles(UfArgs a) {
    Ctl *c = a.ctl;
    const uint32_t nt = ld_acquire(&c->scount);
    if (nt == 0) return;
    const UfHookOp op{a.rows, a.uf, a.roots, c->ufsel[0]};
    expand_tiles<QPL>(op, a.cfg, a.tiles, nt, DirectOut{Sink{nullptr, &c->scount}, TileSink{nullptr, &c->scount}});
}

// Flatten (cc_union_repair.cuh:126-137): each affected vertex takes its
// root as label; the marks of the touched components are cleared (every
// marked label was an affected vertex).
__global__ void k_uf_flatten(UfArgs a) {
    const uint32_t cnt = ld_acquire(&a.ctl->acount);
    if (gthread() == 0) a.ctl->scount = 0;
    for (uint32_t i = gthread(); i < cnt; i += gthreads()) {
        const uint32_t v = a.alist[i];
        const uint32_t root = uf_find(a.uf, v);
        a.state[v] = (uint64_t(root) << 32) | kNone;
        a.uf[v] = root;  // a valid compression, and the invariant uf[v] = label[v]
        a.marks[v >> 5] = 0u;
    }
}

// Inserted edges unite their components.
__global__ void k_uf_union_adds(UfArgs a) {
    for (uint32_t i = gthread(); i < a.na; i += gthreads()) {
        const cig_edge e = a.adds[i];
        if (e.u >= a.n || e.v >= a.n || (a.add_skip && a.add_skip[i])) continue;
        if (uint32_t(a.state[e.u] >> 32) != uint32_t(a.state[e.v] >> 32)) uf_unite(a.uf, e.u, e.v);
    }
}

// After inserts: every vertex whose label's root moved takes the new root
// (uf[v] was its label, so this is at most two steps per vertex).
__global__ void k_uf_relabel(UfArgs a) {
    for (uint32_t v = gthread(); v < a.n; v += gthreads()) {
        const uint32_t label = uint32_t(a.state[v] >> 32);
        const uint32_t root = uf_find(a.uf, label);
        if (root != label) a.state[v] = (uint64_t(root) << 32) | kNone;
        a.uf[v] = root;
    }
}

}  // namespace dev
}  // namespace cig
