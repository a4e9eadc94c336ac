// libcig: MI355X incremental graph engine behind the C-ABI of include/cig.h.
//
// Memory layout in HBM (n vertices):
//   rows[n]        16 B  {pool offset | tier bit, degree, capacity}
//   HBM pool            4-B destination slots; each row owns a power-of-two
//                       block (min 4), rows grow by relocating to the tail
//   pinned host pool    same layout for rows outside the HBM budget; kernels
//                       read them zero-copy over the host link
//   state[n]       8 B  traversal: value<<32 | parent   (SSSP/BFS/CC)
//   rank/comp/res  4 B each (PR, Kahan-compensated rank)
//   q[2][n], stamp[n], affected bitmap, affected list, tiles[2], counters.
// Every batch runs as a chain of kernels on one stream with no host round
// trip; the host reads a 200-B control block at the end of the call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cig.h"
#include "common.h"
#include "host_par.h"
#include "kernels.hpp"
#include "msi.hpp"
#include "ccuf.hpp"

using namespace cig;
using namespace cig::dev;

namespace {

thread_local std::string g_create_error;

// Scratch of cig_query_noalloc: one array per GPU, allocated when the code
// object is loaded on that GPU and held until the process ends. Its name is a
// host-side stand-in, never a device address: host code uses only the
// per-device address that scratch_address() resolves.
static __device__ __align__(256) uint8_t scratch[CIG_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices] = {};

struct HipError : std::runtime_error {
    cig_status code;
    HipError(cig_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(call)                                                                        \
    do {                                                                                    \
        hipError_t _e = (call);                                                             \
        if (_e != hipSuccess)                                                               \
            throw HipError(_e == hipErrorOutOfMemory ? CIG_ENOMEM : CIG_EDEVICE,            \
                           std::string(#call) + ": " + hipGetErrorString(_e));              \
    } while (0)

// This device's address of the scratch array, resolved once per device
// with that device current (every engine's set-up calls this).
uint8_t *scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) throw HipError(CIG_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = static_cast<uint8_t *>(p);
    }
    return g_scratch_ptr[device];
}

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t count = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        count = 0;
    }
    void alloc(size_t n) {
        release();
        const size_t bytes = std::max<size_t>(1, n) * sizeof(T);
        HIPCHK(hipMalloc(&p, bytes));
        count = n;
    }
    size_t bytes() const { return count * sizeof(T); }
};

template <class T>
struct HBuf {  // pinned host memory (hipHostMalloc; never hipHostRegister)
    T *p = nullptr;
    size_t count = 0;
    HBuf() = default;
    HBuf(const HBuf &) = delete;
    HBuf &operator=(const HBuf &) = delete;
    ~HBuf() { release(); }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        count = 0;
    }
    void alloc(size_t n, unsigned flags = hipHostMallocDefault) {
        release();
        HIPCHK(hipHostMalloc(&p, std::max<size_t>(1, n) * sizeof(T), flags));
        count = n;
    }
    T *device_ptr() const {
        void *d = nullptr;
        HIPCHK(hipHostGetDevicePointer(&d, p, 0));
        return static_cast<T *>(d);
    }
};

// A stream and two pinned bounce buffers that move device results into
// caller (pageable) memory: the device copies chunk i into one buffer while
// host threads move chunk i-1 out of the other. The engine's lane uses the
// engine stream; the serving loop has a second lane, so that a query's
// results leave the device while the next update group runs.
struct CopyLane {
    static constexpr uint64_t kChunk = uint64_t(256) << 20;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    HBuf<uint8_t> bounce[2];
    hipEvent_t ev[2] = {};
    CopyLane() = default;
    CopyLane(const CopyLane &) = delete;
    CopyLane &operator=(const CopyLane &) = delete;
    ~CopyLane() { release(); }
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &e : ev) {
            if (e) (void)hipEventDestroy(e);
            e = nullptr;
        }
        for (auto &b : bounce) b.release();
        if (own_stream && stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
        own_stream = false;
    }
    // Uses `s`, or a new stream of its own when s is null.
    void init(hipStream_t s) {
        if (s) {
            stream = s;
        } else {
            HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            own_stream = true;
        }
    }
    // Copies `bytes` from device memory to `out`; returns with the copy complete.
    void copy(void *out, const void *dev, uint64_t bytes) {
        if (bytes <= kChunk) {
            HIPCHK(hipMemcpyAsync(out, dev, bytes, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            return;
        }
        if (!bounce[0].p) {
            for (int b = 0; b < 2; ++b) {
                bounce[b].alloc(kChunk);
                HIPCHK(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
            }
        }
        const unsigned T = pick_threads(0);
        auto host_copy = [&](uint64_t chunk) {
            const uint64_t off = chunk * kChunk, len = std::min(kChunk, bytes - off);
            HIPCHK(hipEventSynchronize(ev[chunk & 1]));
            uint8_t *to = static_cast<uint8_t *>(out) + off;
            const uint8_t *from = bounce[chunk & 1].p;
            parallel_for(len, T, [&](uint64_t lo, uint64_t hi, unsigned) { std::memcpy(to + lo, from + lo, hi - lo); });
        };
        const uint64_t chunks = (bytes + kChunk - 1) / kChunk;
        for (uint64_t i = 0; i < chunks; ++i) {
            const uint64_t off = i * kChunk, len = std::min(kChunk, bytes - off);
            HIPCHK(hipMemcpyAsync(bounce[i & 1].p, static_cast<const uint8_t *>(dev) + off, len,
                                  hipMemcpyDeviceToHost, stream));
            HIPCHK(hipEventRecord(ev[i & 1], stream));
            if (i > 0) host_copy(i - 1);  // frees buffer (i-1)&1 for chunk i+1
        }
        host_copy(chunks - 1);
    }
};

constexpr uint32_t kTileLenDefault = 512;
constexpr uint32_t kSplitDefault = 64;
constexpr uint32_t kAdaptSplitDefault = 1024;
constexpr uint32_t kQplDefault = 2;  // quads (4 slots) per lane and step
// PR rounds whose frontier holds at least n / kPullDivisorDefault vertices
// run in pull form; below n / kPullMaskDivisor the pull tests the
// active-source mask before each gather. (R-MAT scale 25: a masked sweep of
// all incoming rows costs less than scattering with atomics from a frontier
// of n/4096 hub-heavy vertices.)
constexpr uint32_t kPullDivisorDefault = 4096;
constexpr uint32_t kPullMaskDivisor = 4;
constexpr uint32_t kPullBlocksPerCu = 64;

}  // namespace

struct cig_engine {
    cig_config cfg{};
    std::string err;
    cig_server *server = nullptr;  // the serving loop that owns the engine while it runs
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[10] = {};
    // A second compute stream: a batch's invalidation waves run on it while
    // the engine stream groups the batch's records (see traversal_batch).
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // After an overlapped call the side stream clears the call's per-row
    // counts and resets the mutation counters (finish_split): side_done_ev
    // marks the end of that work, which anything else on the engine stream
    // waits for (flush_finish); side_ready says the next overlapped call may
    // group its records without waiting for the engine stream.
    hipEvent_t side_done_ev = nullptr;
    bool side_pending = false, side_ready = false;
    bool fresh_records = false;  // this call enqueued nothing before its batch (resident records)
    bool overlapped = false;  // this call's invalidation ran beside the grouping
    bool scan_timed = false;  // ev[8]/ev[9] bracket this call's repair scan kernel
    // CIG_CFG_KERNEL_TIMING: an event pair per propagation round launch
    static constexpr uint32_t kRoundEvents = 512;
    std::vector<hipEvent_t> rev_begin, rev_end;
    uint32_t round_timed = 0;

    uint32_t n = 0;
    uint32_t nrows = 0;  // n, or 2n with the incoming rows [n, 2n)
    bool rev = false;    // incoming-row index kept (pull repair)
    bool sym = false;    // symmetric graph: the outgoing rows are the incoming rows
    bool loaded = false, initialized = false;
    uint32_t source = 0;
    // PR vertex renumbering (see degree_order): perm[v] = internal id of the
    // caller's vertex v, inv its inverse; empty when ids are kept
    bool renum = false;
    std::vector<uint32_t> perm_h, inv_h;
    DBuf<uint32_t> perm_d;

    DBuf<Row> rows;
    DBuf<uint64_t> state;
    DBuf<float> rank, comp, res, delta;
    DBuf<uint32_t> q0, q1, stamp, aff, alist, filt, hot, hist, score;
    DBuf<Tile> t0, t1;
    uint64_t tile_cap = 0;
    // PR pull plan: tiles of long incoming rows, rows of several tiles, partial sums
    DBuf<Tile> ptiles;
    DBuf<PullRow> prow;
    DBuf<float> tpart;
    DBuf<unsigned long long> fl_off;  // host-tier free blocks by class (FreeList), when the tier holds rows
    DBuf<int> fl_cnt;
    static constexpr uint32_t kFlCap = 1u << 20;  // blocks listed per class
    FreeList free_list() const { return FreeList{fl_off.p, fl_cnt.p, fl_off.p ? kFlCap : 0u}; }
    void clear_free_list() {
        if (fl_cnt.p) fill32(reinterpret_cast<uint32_t *>(fl_cnt.p), kFlClasses, 0u);
    }
    DBuf<unsigned long long> amask;  // active-source mask of the pull passes
    DBuf<uint32_t> odeg;             // out-degrees for the dense consume of one solve
    DBuf<unsigned long long> adm_hist, adm_acc;  // cache admission: capacity per score, sums
    uint64_t ptile_cap = 0;
    DBuf<uint64_t> cnt64, gofs;
    DBuf<uint32_t> rslot;  // per record of a call: rank within its row's deletes / adds
    DBuf<uint32_t> touched, grouped, small_list, mid_list, mid16_list, big_list, add_list, overflow;
    DBuf<AddPiece> pieces;              // insert-phase copy pieces
    DBuf<StatShard> stats;              // sharded statistics (see k_reduce_stats)
    DBuf<uint32_t> htoff, hkeys, hcnt;  // long-row delete tables
    DBuf<uint8_t> dclass;               // per-row delete class of the current call
    // grouped batches: (u,v) table, op lists, cancel flags, per-batch cuts
    DBuf<unsigned long long> gkey;
    DBuf<uint32_t> ghead, gnext, gcut, gover;
    DBuf<uint8_t> gconf, gskip;
    const uint8_t *cur_skip = nullptr;  // cancel flags of the group being applied
    // multi-source queries
    DBuf<uint32_t> ms_dist, ms_stamp, ms_out;
    uint8_t *scratch_dev = nullptr;  // this device's copy of the no-allocation scratch array
    // incremental multi-source state (cig_init with 2..64 BFS/SSSP sources)
    uint32_t msk = 0;  // sources kept; 0: the single-source state word
    std::vector<uint32_t> ms_sources;
    DBuf<unsigned long long> msi_st, msi_act, msi_snap, msi_aff, msi_pinv;
    // CC by union-find (CIG_CFG_CC_UNION_REPAIR)
    DBuf<uint32_t> uf, uroot;
    DBuf<unsigned long long> ms_act, ms_snap, ms_seen;
    CopyLane lane;  // large device-to-host results on the engine stream
    DBuf<uint2> hcnts;
    DBuf<DelTile> dtiles;
    uint64_t rec_cap = 0;
    DBuf<cig_edge> rec_dels, rec_adds;
    DBuf<Ctl> ctl;
    HBuf<Ctl> hctl;
    HBuf<uint64_t> prog;  // round progress word written by the GPU (coherent, mapped)
    uint64_t *prog_dev = nullptr;
    uint64_t chain_epoch = 0;
    HBuf<CtlPub> pub;     // the control block as a call's last kernel publishes it (coherent, mapped)
    CtlPub *pub_dev = nullptr;
    uint64_t pub_seq = 0;

    DBuf<uint32_t> hbm_pool;
    HBuf<uint32_t> host_pool;
    uint32_t *host_pool_dev = nullptr;
    uint64_t hbm_slots = 0, host_slots = 0;
    uint64_t budget_slots = 0;  // 0 = unlimited

    // resident update stream
    DBuf<cig_edge> s_adds, s_dels;
    std::vector<uint32_t> batch_sizes;
    uint32_t next_batch = 0;
    uint64_t add_off = 0, del_off = 0;

    // launch shape
    uint32_t split = kSplitDefault, tile_len = kTileLenDefault, qpl = kQplDefault;
    // frontier rounds keep rows of up to this many slots whole when waves
    // take few items (ExpandCfg::adapt); 0 = always cut at `split`
    uint32_t adapt_split = kAdaptSplitDefault;
    uint32_t rblock = kBlock;  // relax round workgroup size (256 or 512)
    uint32_t grid_small = 0;   // one-shot kernels
    uint32_t grid_coop = 0;    // persistent cooperative kernels
    uint32_t grid_round = 0;   // chained round kernels
    uint32_t grid_pull = 0;    // PR pull-form kernels (whole-graph sweeps)
    bool coop = true;

    // ------------------------------------------------------------ plumbing
    Pools pools() const { return Pools{hbm_pool.p, host_pool_dev}; }
    // Workgroups of 256 threads for `items` work items, `per` per thread, at
    // most grid_small and at least one.
    uint32_t grid_for(uint64_t items, uint32_t per) const {
        const uint64_t g = (items + 256ull * per - 1) / (256ull * per);
        return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(grid_small, g)));
    }
    ExpandCfg ecfg(uint32_t stage) const { return ExpandCfg{pools(), split, tile_len, stage, stats.p}; }
    uint32_t weight_kind() const {
        switch (cfg.algo) {
        case CIG_ALGO_BFS: return 2;
        case CIG_ALGO_CC: return 3;
        default: return cfg.weight_mode == CIG_WEIGHT_HASH64 ? 1u : 0u;
        }
    }
    bool is_pr() const { return cfg.algo == CIG_ALGO_PR; }

    void setup_device() {
        HIPCHK(hipSetDevice(cfg.device));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, cfg.device));
        cus = prop.multiProcessorCount;
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&side_done_ev, hipEventDisableTiming));
        scratch_dev = scratch_address(cfg.device);
        lane.init(stream);
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
        if (cfg.flags & CIG_CFG_KERNEL_TIMING) {
            rev_begin.resize(kRoundEvents);
            rev_end.resize(kRoundEvents);
            for (uint32_t i = 0; i < kRoundEvents; ++i) {
                HIPCHK(hipEventCreate(&rev_begin[i]));
                HIPCHK(hipEventCreate(&rev_end[i]));
            }
        }
        ctl.alloc(1);
        hctl.alloc(1);
        prog.alloc(1, hipHostMallocCoherent | hipHostMallocMapped);
        *prog.p = 0;
        prog_dev = prog.device_ptr();
        pub.alloc(1, hipHostMallocCoherent | hipHostMallocMapped);
        std::memset(pub.p, 0, sizeof(CtlPub));
        pub_dev = pub.device_ptr();
        stats.alloc(kStatShards);
        HIPCHK(hipMemsetAsync(stats.p, 0, stats.bytes(), stream));
        std::memset(hctl.p, 0, sizeof(Ctl));
        HIPCHK(hipMemcpyAsync(ctl.p, hctl.p, sizeof(Ctl), hipMemcpyHostToDevice, stream));
        if (cfg.split_degree) split = std::max<uint32_t>(64, cfg.split_degree);
        if (cfg.tile_len) tile_len = (std::max<uint32_t>(64, cfg.tile_len) + 3u) & ~3u;  // whole quads
        if (cfg.adapt_split) adapt_split = cfg.adapt_split <= split ? 0u : cfg.adapt_split;
        // edges per lane come in whole quads: 4, 8 or 16
        if (cfg.edges_per_lane) qpl = cfg.edges_per_lane >= 16 ? 4 : cfg.edges_per_lane >= 8 ? 2 : 1;
        coop = (cfg.flags & CIG_CFG_COOPERATIVE) != 0;
        if (coop && !prop.cooperativeLaunch) throw HipError(CIG_EDEVICE, "device lacks cooperative launch");
        grid_small = uint32_t(cus) * 8;
        grid_pull = uint32_t(cus) * (cfg.pull_blocks_per_cu ? std::min<uint32_t>(cfg.pull_blocks_per_cu, 128) : kPullBlocksPerCu);
        grid_round = uint32_t(cus) * (cfg.round_blocks_per_cu ? std::min<uint32_t>(cfg.round_blocks_per_cu, 16) : 1);
    }

    void sync() { HIPCHK(hipStreamSynchronize(stream)); }

    template <class K>
    uint32_t coop_grid_for(K kernel) {
        int per_cu = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(kernel), kBlock, 0));
        per_cu = std::max(1, std::min(per_cu, 4));
        return uint32_t(per_cu) * uint32_t(cus);
    }

    template <class K, class A>
    void launch_coop(K kernel, A &args, uint32_t grid) {
        void *kargs[] = {&args};
        HIPCHK(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(kernel), dim3(grid), dim3(kBlock),
                                          kargs, 0, stream));
    }

    void fill32(uint32_t *p, uint64_t count, uint32_t value) {
        if (!count) return;
        k_fill32<<<grid_small, 256, 0, stream>>>(p, count, value);
        HIPCHK(hipGetLastError());
    }

    // Incoming rows in CSR form: the sources of each vertex's in-edges,
    // ascending. Threads take source ranges of about m/T edges and bucket
    // (dst, src) pairs by destination block; each block is then counting-
    // sorted by destination, which keeps the sources in ascending order.
    void build_incoming(const uint64_t *off, const uint32_t *dst, unsigned T, std::vector<uint64_t> &ioff,
                        std::vector<uint32_t> &isrc) {
        const uint64_t m = off[n];
        ioff.assign(size_t(n) + 1, 0);
        isrc.resize(m);
        unsigned bb = 0;  // destination block = 2^bb vertices, at most 4096 blocks
        while ((uint64_t(n) >> bb) > 4096) ++bb;
        const uint64_t B = ((uint64_t(n) - 1) >> bb) + 1;
        std::vector<uint32_t> sb(T + 1, n);
        for (unsigned t = 0; t < T; ++t)
            sb[t] = uint32_t(std::lower_bound(off, off + n, m * t / T) - off);
        std::vector<uint64_t> cnt(size_t(T) * B, 0);
        std::vector<std::thread> pool;
        auto each_thread = [&](auto f) {
            pool.clear();
            for (unsigned t = 0; t < T; ++t) pool.emplace_back([&, t] { f(t); });
            for (auto &th : pool) th.join();
        };
        each_thread([&](unsigned t) {
            uint64_t *c = cnt.data() + size_t(t) * B;
            for (uint64_t e = off[sb[t]]; e < off[sb[t + 1]]; ++e) c[dst[e] >> bb]++;
        });
        std::vector<uint64_t> bstart(B + 1, 0);
        uint64_t run = 0;
        for (uint64_t b = 0; b < B; ++b) {
            bstart[b] = run;
            for (unsigned t = 0; t < T; ++t) {
                const uint64_t x = cnt[size_t(t) * B + b];
                cnt[size_t(t) * B + b] = run;
                run += x;
            }
        }
        bstart[B] = m;
        std::vector<uint64_t> pairs(m);
        each_thread([&](unsigned t) {
            uint64_t *c = cnt.data() + size_t(t) * B;
            for (uint32_t s = sb[t]; s < sb[t + 1]; ++s)
                for (uint64_t e = off[s]; e < off[s + 1]; ++e) pairs[c[dst[e] >> bb]++] = uint64_t(dst[e]) << 32 | s;
        });
        std::atomic<uint64_t> next{0};
        each_thread([&](unsigned) {
            std::vector<uint64_t> local(size_t(1) << bb);
            for (uint64_t b; (b = next.fetch_add(1)) < B;) {
                const uint64_t base = b << bb, nb = std::min<uint64_t>(uint64_t(1) << bb, n - base);
                std::fill(local.begin(), local.begin() + int64_t(nb), 0);
                for (uint64_t i = bstart[b]; i < bstart[b + 1]; ++i) local[(pairs[i] >> 32) - base]++;
                uint64_t r = bstart[b];
                for (uint64_t k = 0; k < nb; ++k) {
                    ioff[base + k] = r;
                    const uint64_t x = local[k];
                    local[k] = r;
                    r += x;
                }
                for (uint64_t i = bstart[b]; i < bstart[b + 1]; ++i)
                    isrc[local[(pairs[i] >> 32) - base]++] = uint32_t(pairs[i]);
            }
        });
        ioff[n] = m;
    }

    // Budget placement order: decreasing capacity (a power of two), outgoing
    // before incoming rows of the same capacity, row id ascending. Long rows
    // of either direction come first: besides traversal reads, a batch's
    // deletes scan every long row they touch, and the hub rows are touched
    // by nearly every batch.
    static std::vector<uint32_t> placement_order(const std::vector<Row> &hrows, uint32_t n) {
        auto key = [&](uint32_t v) -> uint32_t {
            const uint32_t cap = hrows[v].cap;
            const uint32_t lg = cap ? 31u - uint32_t(__builtin_clz(cap)) : 0u;
            return 2u * (cap ? 32u - lg : 33u) + (v < n ? 0u : 1u);
        };
        const uint32_t nrows = uint32_t(hrows.size());
        std::vector<uint64_t> start(69, 0);  // keys 2..67
        for (uint32_t v = 0; v < nrows; ++v) start[key(v) + 1]++;
        for (int k = 0; k < 68; ++k) start[k + 1] += start[k];
        std::vector<uint32_t> order(nrows);
        for (uint32_t v = 0; v < nrows; ++v) order[start[key(v)]++] = v;
        return order;
    }

    uint32_t round_bound() const { return uint32_t(std::min<uint64_t>(2ull * n + 64, 0x7FFFFFFFull)); }

    // Incoming rows for the pull repair and PageRank's gathers: rows [n, 2n),
    // or the outgoing rows themselves when the graph is symmetric.
    bool pull_rows() const { return rev || sym; }
    const Row *in_rows_p() const { return rev ? rows.p + n : rows.p; }

    // CIG_CFG_SYMMETRIC: every vertex's in-degree must equal its out-degree
    // (necessary for a symmetric multiset of edges; the caller asserts the rest).
    static void check_symmetric_degrees(uint32_t nv, const uint64_t *off, const uint32_t *dst, unsigned T) {
        std::vector<uint32_t> indeg(nv, 0u);
        parallel_for(off[nv], T, [&](uint64_t lo, uint64_t hi, unsigned) {
            for (uint64_t e = lo; e < hi; ++e) __atomic_fetch_add(&indeg[dst[e]], 1u, __ATOMIC_RELAXED);
        });
        std::atomic<int> bad{0};
        parallel_for(nv, T, [&](uint64_t lo, uint64_t hi, unsigned) {
            for (uint64_t v = lo; v < hi; ++v)
                if (uint64_t(indeg[v]) != off[v + 1] - off[v]) bad = 1;
        });
        if (bad) throw HipError(CIG_EINVAL, "CIG_CFG_SYMMETRIC: a vertex's in-degree differs from its out-degree");
    }

    // PageRank's dense rounds gather one contribution per incoming edge, from
    // the edge's source. Internal ids by decreasing out-degree (ties by id;
    // degrees of 4095 and more form the first class) put the contributions of
    // the hubs, which most gathers read, in a few megabytes at the front of
    // the per-vertex arrays (R-MAT scale 25: the first 2^20 internal ids are
    // the sources of 80% of the edges), where they share cache lines and stay
    // in L2. A stable counting sort over contiguous thread chunks.
    static void degree_order(uint32_t nv, const uint64_t *off, unsigned T, std::vector<uint32_t> &perm,
                             std::vector<uint32_t> &inv) {
        constexpr uint32_t K = 4096;
        perm.assign(nv, 0u);
        inv.assign(nv, 0u);
        auto key = [&](uint64_t v) { return uint32_t(K - 1 - std::min<uint64_t>(off[v + 1] - off[v], K - 1)); };
        std::vector<uint64_t> cnt(size_t(T) * K, 0);
        parallel_for(nv, T, [&](uint64_t lo, uint64_t hi, unsigned t) {
            uint64_t *c = cnt.data() + size_t(t) * K;
            for (uint64_t v = lo; v < hi; ++v) c[key(v)]++;
        });
        uint64_t run = 0;
        for (uint32_t k = 0; k < K; ++k)
            for (unsigned t = 0; t < T; ++t) {
                const uint64_t x = cnt[size_t(t) * K + k];
                cnt[size_t(t) * K + k] = run;
                run += x;
            }
        parallel_for(nv, T, [&](uint64_t lo, uint64_t hi, unsigned t) {
            uint64_t *c = cnt.data() + size_t(t) * K;
            for (uint64_t v = lo; v < hi; ++v) {
                const uint32_t at = uint32_t(c[key(v)]++);
                inv[at] = uint32_t(v);
                perm[v] = at;
            }
        });
    }

    // The graph in internal ids: row i is the caller's row inv[i], its
    // destinations mapped through perm.
    static void permute_csr(uint32_t nv, const uint64_t *off, const uint32_t *dst, unsigned T,
                            const std::vector<uint32_t> &perm, const std::vector<uint32_t> &inv,
                            std::vector<uint64_t> &poff, std::vector<uint32_t> &pdst) {
        poff.assign(size_t(nv) + 1, 0);
        std::vector<uint64_t> part(T + 1, 0);
        parallel_for(nv, T, [&](uint64_t lo, uint64_t hi, unsigned t) {
            uint64_t s = 0;
            for (uint64_t i = lo; i < hi; ++i) s += off[inv[i] + 1] - off[inv[i]];
            part[t + 1] = s;
        });
        for (unsigned t = 0; t < T; ++t) part[t + 1] += part[t];
        parallel_for(nv, T, [&](uint64_t lo, uint64_t hi, unsigned t) {
            uint64_t o = part[t];
            for (uint64_t i = lo; i < hi; ++i) {
                poff[i] = o;
                o += off[inv[i] + 1] - off[inv[i]];
            }
        });
        poff[nv] = off[nv];
        pdst.resize(off[nv]);
        parallel_for(nv, T, [&](uint64_t lo, uint64_t hi, unsigned) {
            for (uint64_t i = lo; i < hi; ++i) {
                const uint32_t *s = dst + off[inv[i]];
                uint32_t *d = pdst.data() + poff[i];
                for (uint64_t k = 0, deg = poff[i + 1] - poff[i]; k < deg; ++k) d[k] = perm[s[k]];
            }
        });
    }

    // Caller ids of update records to internal ids, in place (ids >= n stay,
    // so invalid records are rejected as before).
    void records_to_internal(cig_edge *p, uint64_t cnt) {
        if (!renum || cnt == 0) return;
        k_map_records<<<grid_small, 256, 0, stream>>>(p, cnt, perm_d.p, n);
        HIPCHK(hipGetLastError());
    }

    // ------------------------------------------------------------ load
    void load(uint32_t nv, const uint64_t *off, const uint32_t *dst) {
        if (nv == 0) throw HipError(CIG_EINVAL, "graph must have at least one vertex");
        finish_pending = false;  // the buffers it names are about to be replaced
        const unsigned T = pick_threads(0);
        std::atomic<int> bad{0};
        parallel_for(nv, T, [&](uint64_t lo, uint64_t hi, unsigned) {
            for (uint64_t v = lo; v < hi; ++v)
                if (off[v + 1] < off[v] || off[v + 1] - off[v] > 0xFFFFFFFFull) bad = 1;
        });
        if (bad) throw HipError(CIG_EINVAL, "offsets must be non-decreasing, row length < 2^32");
        const uint64_t m = off[nv];
        parallel_for(m, T, [&](uint64_t lo, uint64_t hi, unsigned) {
            for (uint64_t e = lo; e < hi; ++e)
                if (dst[e] >= nv) bad = 1;
        });
        if (bad) throw HipError(CIG_EINVAL, "destination id out of range");
        renum = is_pr() && !(cfg.flags & CIG_CFG_PR_KEEP_IDS);
        std::vector<uint64_t> poff;
        std::vector<uint32_t> pdst;
        if (renum) {
            degree_order(nv, off, T, perm_h, inv_h);
            permute_csr(nv, off, dst, T, perm_h, inv_h, poff, pdst);
            off = poff.data();
            dst = pdst.data();
        } else {
            perm_h.clear();
            inv_h.clear();
        }
        n = nv;
        budget_slots = cfg.cache_bytes / sizeof(uint32_t);
        // incoming rows (the device reverse index) for deletion repair
        // (PR: the pull rounds' gather rows; the cooperative driver only scatters)
        const bool pull = !(cfg.flags & CIG_CFG_SCAN_REPAIR) && !(is_pr() && ((cfg.flags & CIG_CFG_PR_PUSH_ONLY) || coop));
        sym = pull && (cfg.flags & CIG_CFG_SYMMETRIC);
        rev = pull && !sym;
        if (sym) check_symmetric_degrees(nv, off, dst, T);
        if (rev && n > 0x7FFFFFFFu) throw HipError(CIG_ELIMIT, "incoming index needs n < 2^31 (use CIG_CFG_SCAN_REPAIR)");
        nrows = rev ? 2 * n : n;
        std::vector<uint64_t> ioff;
        std::vector<uint32_t> isrc;
        if (rev) build_incoming(off, dst, T, ioff, isrc);
        auto row_data = [&](uint32_t v) -> const uint32_t * {
            return v < n ? dst + off[v] : isrc.data() + ioff[v - n];
        };

        std::vector<Row> hrows(nrows);
        std::vector<uint64_t> part(T, 0);
        parallel_for(nrows, T, [&](uint64_t lo, uint64_t hi, unsigned t) {
            uint64_t sum = 0;
            for (uint64_t v = lo; v < hi; ++v) {
                hrows[v].deg = uint32_t(v < n ? off[v + 1] - off[v] : ioff[v - n + 1] - ioff[v - n]);
                hrows[v].cap = uint32_t(next_pow2_cap(hrows[v].deg));
                sum += hrows[v].cap;
            }
            part[t] = sum;
        });
        const uint64_t total_cap = std::accumulate(part.begin(), part.end(), uint64_t(0));
        // tier placement: rows by decreasing capacity while the budget lasts
        // (90% fill); incoming rows are read by the deletion repair, by
        // PageRank's pull rounds and by the deletes' long-row scans
        std::vector<uint8_t> in_hbm(nrows, 1);
        uint64_t hbm_need = total_cap, host_need = 0;
        if (budget_slots && total_cap > budget_slots * 9 / 10) {
            const std::vector<uint32_t> order = placement_order(hrows, n);
            uint64_t used = 0;
            const uint64_t limit = budget_slots * 9 / 10;
            for (uint32_t v : order) {
                const bool fits = hrows[v].deg != 0 && used + hrows[v].cap <= limit;
                in_hbm[v] = fits ? 1 : 0;
                if (fits) used += hrows[v].cap;
            }
            hbm_need = used;
            host_need = total_cap - used;
        }
        const uint64_t slack = cfg.pool_slack ? cfg.pool_slack : (uint64_t(1) << 22);
        hbm_slots = budget_slots ? budget_slots : hbm_need + hbm_need / 4 + slack;
        // the host tier takes the budget's evictions and the growth of the cold
        // rows, and packing it costs seconds at scale 27: half its rows' capacity
        // again as headroom (pinned host memory is the cheaper resource)
        host_slots = host_need ? host_need + host_need / 2 + std::min<uint64_t>(slack, uint64_t(1) << 20)
                               : (uint64_t(1) << 16);
        hbm_pool.alloc(hbm_slots);
        host_pool.alloc(host_slots);
        host_pool_dev = host_pool.device_ptr();
        if (host_need) {
            if (!fl_off.p) {
                fl_off.alloc(uint64_t(kFlClasses) * kFlCap);
                fl_cnt.alloc(kFlClasses);
            }
        } else {
            fl_off.release();
            fl_cnt.release();
        }
        clear_free_list();
        // rows keep row-id order within each tier; the copies run in parallel
        uint64_t htail = 0, ctail = 0;
        for (uint32_t v = 0; v < nrows; ++v) {
            Row &r = hrows[v];
            if (in_hbm[v]) {
                r.off = htail;
                htail += r.cap;
            } else {
                r.off = ctail | kHostTier;
                ctail += r.cap;
            }
        }
        std::vector<uint32_t> stage(htail, 0u);
        parallel_for(nrows, T, [&](uint64_t lo, uint64_t hi, unsigned) {
            for (uint64_t v = lo; v < hi; ++v) {
                const Row &r = hrows[v];
                if (!r.deg) continue;
                uint32_t *to = (r.off & kHostTier) ? host_pool.p + (r.off & ~kHostTier) : stage.data() + r.off;
                std::memcpy(to, row_data(uint32_t(v)), size_t(r.deg) * sizeof(uint32_t));
            }
        });
        if (!stage.empty())
            HIPCHK(hipMemcpyAsync(hbm_pool.p, stage.data(), stage.size() * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, stream));
        rows.alloc(nrows);
        HIPCHK(hipMemcpyAsync(rows.p, hrows.data(), size_t(nrows) * sizeof(Row), hipMemcpyHostToDevice, stream));
        sync();  // the staging vectors go out of scope below

        // per-vertex arrays
        q0.alloc(n);
        q1.alloc(n);
        stamp.alloc(n);
        // activity per row: expansions of outgoing rows, records of any row
        hot.alloc(nrows);
        hist.alloc(nrows);
        score.alloc(nrows);
        // per-row mutation arrays (outgoing and incoming rows)
        cnt64.alloc(nrows);
        gofs.alloc(nrows);
        touched.alloc(nrows);
        small_list.alloc(nrows);
        mid_list.alloc(nrows);
        mid16_list.alloc(nrows);
        big_list.alloc(nrows);
        add_list.alloc(nrows);
        overflow.alloc(nrows);
        htoff.alloc(nrows);
        hcnt.alloc(nrows);
        dclass.alloc(nrows);
        if (is_pr()) {
            rank.alloc(n);
            comp.alloc(n);
            res.alloc(n);
            delta.alloc(n);
        } else {
            state.alloc(n);
            aff.alloc((uint64_t(n) + 31) / 32);
            alist.alloc(n);
            filt.alloc(kFilterWords);
        }
        ensure_tiles();
        fill32(stamp.p, n, 0);
        fill32(hot.p, nrows, 0);
        fill32(hist.p, nrows, 0);
        fill32(score.p, nrows, 0);
        fill32(hcnt.p, nrows, 0);
        fill32(reinterpret_cast<uint32_t *>(cnt64.p), uint64_t(nrows) * 2, 0);
        if (!is_pr()) fill32(aff.p, (uint64_t(n) + 31) / 32, 0);
        // control block: pools
        std::memset(hctl.p, 0, sizeof(Ctl));
        hctl.p->round = 1;
        hctl.p->pool_tail[0] = htail;
        hctl.p->pool_tail[1] = ctail;
        hctl.p->pool_cap[0] = hbm_slots;
        hctl.p->pool_cap[1] = host_slots;
        HIPCHK(hipMemcpyAsync(ctl.p, hctl.p, sizeof(Ctl), hipMemcpyHostToDevice, stream));
        if (renum) {
            perm_d.alloc(n);
            HIPCHK(hipMemcpyAsync(perm_d.p, perm_h.data(), size_t(n) * 4, hipMemcpyHostToDevice, stream));
        }
        sync();
        loaded = true;
        initialized = false;
    }

    // Tile lists hold at most one round's worth of tiles: every row of the
    // round at most once, so sum(deg)/tile_len + rows >= split.
    void ensure_tiles() {
        const uint64_t slots = hbm_slots + host_slots;
        const uint64_t need = slots / tile_len + slots / split + 1024;
        if (need <= tile_cap) return;
        t0.alloc(need);
        t1.alloc(need);
        tile_cap = need;
    }

    // Per-record arrays of a call: the grouped destinations and the ranks.
    // Per-call buffers grow to at least half again what they held, so calls of
    // growing size (a serving loop's groups) reallocate O(log) times: every
    // reallocation frees the old buffer, and the free waits for the device.
    template <class T>
    bool grow(DBuf<T> &b, uint64_t need) {
        if (b.count >= need) return false;
        retire(b);
        b.alloc(std::max<uint64_t>(need, b.count + b.count / 2));
        return true;
    }

    // A buffer replaced by a larger one is freed with the engine, not at once:
    // hipFree waits for the whole device, which would hold this call behind
    // work on the engine's other streams (a serving loop's result copies).
    // Growth is geometric, so what is kept stays below the buffers' own size.
    std::vector<void *> retired;
    template <class T>
    void retire(DBuf<T> &b) {
        if (b.p) retired.push_back(b.p);
        b.p = nullptr;  // count kept: the next size grows from it
    }

    void ensure_grouped(uint64_t records) {
        grow(grouped, records);
        grow(rslot, records);
    }

    void ensure_records(uint64_t nd, uint64_t na) {
        const uint64_t need = std::max(nd, na);
        grow(rec_dels, need);
        grow(rec_adds, need);
        ensure_grouped(nd + na);
    }

    // ------------------------------------------------------------ control
    void read_ctl() {
        HIPCHK(hipMemcpyAsync(hctl.p, ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, stream));
        sync();
    }

    void reset_ctl(bool keep_rounds, bool keep_mut = false) {
        k_reset_ctl<<<1, 64, 0, stream>>>(ctl.p, (keep_rounds ? kResetKeepRounds : 0u) | (keep_mut ? kResetKeepMut : 0u));
        HIPCHK(hipGetLastError());
    }

    LoopArgs loop_args(uint32_t max_rounds) {
        LoopArgs a{};
        a.ctl = ctl.p;
        a.rows = rows.p;
        a.state = state.p;
        a.stamp = stamp.p;
        // activity only feeds cache admission: no budget, no counting
        a.hot = (cfg.flags & CIG_CFG_NO_HOTNESS) || !budget_slots ? nullptr : hot.p;
        a.q0 = q0.p;
        a.q1 = q1.p;
        a.t0 = t0.p;
        a.t1 = t1.p;
        a.cfg = ecfg(kStagePropagate);
        a.cfg.adapt = adapt_split;
        a.seed = cfg.weight_seed;
        a.max_rounds = max_rounds;
        a.delta = ordered() ? (cfg.repair_delta ? cfg.repair_delta : 128u) : 0u;
        return a;
    }

    // Delta-ordered relaxation rounds: SSSP with chained rounds only.
    bool ordered() const {
        return (cfg.flags & CIG_CFG_ORDERED_REPAIR) && cfg.algo == CIG_ALGO_SSSP && !coop;
    }

    // Enqueues round kernels back to back in growing chunks (8, 16, ... 128)
    // and reads the control block once per chunk; `more` says whether the
    // fixed point still has work. Rounds past the bound raise an error.
    // Runs one chain of round kernels. launch(tag, seq) enqueues round seq
    // (1-based); every round posts one progress word at its start (see
    // post_progress) saying whether it has work. Rounds are launched
    // kLookahead ahead of the last word seen, so the host never blocks on
    // the stream and at most kLookahead no-op rounds follow the fixed point.
    // Returns the number of rounds that had work.
    template <class Launch>
    uint64_t run_chain(Launch launch, uint64_t bound, const char *what, hipStream_t on = nullptr) {
        if (!on) on = stream;
        constexpr uint64_t kLookahead = 3;
        const uint64_t epoch = ++chain_epoch & 0xFFFFFFFFull;
        uint64_t issued = 0, seen = 0;
        for (;;) {
            while (issued < bound && issued < seen + kLookahead) {
                ++issued;
                launch((epoch << 31) | issued, uint32_t(issued));
            }
            HIPCHK(hipGetLastError());
            uint64_t w = 0;
            for (uint32_t spin = 0;; ++spin) {
                w = __atomic_load_n(prog.p, __ATOMIC_ACQUIRE);
                if ((w >> 32) == epoch && ((w >> 1) & 0x7FFFFFFFull) > seen) break;
                if ((spin & 1023) == 1023) {
                    const hipError_t q = hipStreamQuery(on);
                    if (q != hipSuccess && q != hipErrorNotReady)
                        throw HipError(CIG_EDEVICE, std::string(what) + ": " + hipGetErrorString(q));
                    if (q == hipSuccess) {  // stream idle: the word must be final now
                        w = __atomic_load_n(prog.p, __ATOMIC_ACQUIRE);
                        if ((w >> 32) == epoch && ((w >> 1) & 0x7FFFFFFFull) > seen) break;
                        throw HipError(CIG_EINTERNAL, std::string(what) + ": round progress lost");
                    }
                }
                __builtin_ia32_pause();
            }
            seen = (w >> 1) & 0x7FFFFFFFull;
            if (!(w & 1ull)) return seen - 1;  // round `seen` found nothing to do
            if (seen >= bound) {
                if (cfg.max_rounds || is_pr()) return seen;  // a configured cap
                throw HipError(CIG_EINTERNAL, std::string(what) + " did not converge within the round bound");
            }
        }
    }

    uint64_t round_launches = 0;  // timed round launches of the current call
    uint64_t inval_waves = 0;     // invalidation waves with work in the current call
    uint64_t calls = 0;           // calls begun (selects the timed ones)
    bool timing_now = false;      // this call's round kernels are timed
    bool stages_now = true;       // this call records its stage markers

    // Stage marker k (ev[k]) of a sampled call. Each event record costs the
    // stream a few microseconds of idle, so calls that kernel_timing_every
    // skips record only their begin/end pair.
    void mark(int k) {
        if (stages_now) HIPCHK(hipEventRecord(ev[k], stream));
    }

    // Launches one propagation round kernel, bracketed by an event pair when
    // kernel timing is on for this call.
    template <class F>
    void timed_round(F launch) {
        const bool timed = timing_now && round_timed < kRoundEvents;
        if (timed) HIPCHK(hipEventRecord(rev_begin[round_timed], stream));
        launch();
        if (timed) {
            HIPCHK(hipEventRecord(rev_end[round_timed++], stream));
            ++round_launches;
        }
    }

    void advance_round(uint64_t by) {
        if (!by) return;
        k_advance_round<<<1, 64, 0, stream>>>(ctl.p, uint32_t(by));
        HIPCHK(hipGetLastError());
    }

    static bool relax_pending(const Ctl &c) {
        return c.qcount[c.round % 3] || c.tcount[c.round % 3];
    }

    template <int W, bool S, int QPL>
    void relax_impl(uint32_t launches_cap) {
        if (coop) {
            auto kernel = k_relax_loop<W, S, QPL>;
            static thread_local uint32_t grid = 0;
            if (!grid) grid = coop_grid_for(kernel);
            grid_coop = grid;
            // Without a configured cap the bound is 2n+64 rounds: a frontier
            // round finalises every vertex one more hop away (one more for a
            // deferred tile), so a larger count means a broken invariant.
            const uint32_t cap = cfg.max_rounds ? cfg.max_rounds : round_bound();
            LoopArgs a = loop_args(cap);
            launch_coop(kernel, a, grid);
            if (cfg.max_rounds) {
                for (uint32_t i = 0; i < launches_cap; ++i) {
                    read_ctl();
                    if (!relax_pending(*hctl.p)) break;
                    launch_coop(kernel, a, grid);
                }
            }
        } else if (ordered()) {
            LoopArgs a = loop_args(1);
            const uint64_t bound = cfg.max_rounds ? cfg.max_rounds : round_bound();
            k_ord_reset<<<1, 64, 0, stream>>>(ctl.p);
            advance_round(run_chain(
                [&](uint64_t tag, uint32_t seq) {
                    k_ord_min<<<grid_round, kBlock, 0, stream>>>(a, seq);
                    timed_round([&] {
                        k_relax_round<W, S, QPL, kBlock, true><<<grid_round, kBlock, 0, stream>>>(a, prog_dev, tag, seq);
                    });
                },
                bound, "ordered relax fixed point"));
        } else {
            LoopArgs a = loop_args(1);
            const uint64_t bound = cfg.max_rounds ? cfg.max_rounds : round_bound();
            advance_round(run_chain(
                [&](uint64_t tag, uint32_t seq) {
                    timed_round([&] {
                        if (rblock == 512)
                            k_relax_round<W, S, QPL, 512><<<grid_round, 512, 0, stream>>>(a, prog_dev, tag, seq);
                        else
                            k_relax_round<W, S, QPL><<<grid_round, kBlock, 0, stream>>>(a, prog_dev, tag, seq);
                    });
                },
                bound, "relax fixed point"));
        }
        HIPCHK(hipGetLastError());
    }

    template <int QPL>
    void relax_dispatch_qpl() {
        switch (weight_kind()) {
        case 0: relax_impl<0, false, QPL>(1u << 20); break;
        case 1: relax_impl<1, false, QPL>(1u << 20); break;
        case 2: relax_impl<2, false, QPL>(1u << 20); break;
        default: relax_impl<3, true, QPL>(1u << 20); break;
        }
    }

    void relax() {
        switch (qpl) {
        case 1: relax_dispatch_qpl<1>(); break;
        case 4: relax_dispatch_qpl<4>(); break;
        default: relax_dispatch_qpl<2>(); break;
        }
    }

    PrArgs pr_args(uint32_t max_rounds) {
        PrArgs a{};
        a.ctl = ctl.p;
        a.rows = rows.p;
        a.rank = rank.p;
        a.comp = comp.p;
        a.res = res.p;
        a.delta = delta.p;
        a.queued = stamp.p;
        a.q0 = q0.p;
        a.q1 = q1.p;
        a.t0 = t0.p;
        a.t1 = t1.p;
        a.cfg = ecfg(kStagePropagate);
        a.eps = cfg.pr_epsilon;
        a.max_rounds = max_rounds;
        a.n = n;
        a.pull_min = kNone;
        if (pull_rows() && !coop) {
            const uint32_t div = cfg.pr_pull_divisor ? cfg.pr_pull_divisor : kPullDivisorDefault;
            a.pull_min = std::max<uint32_t>(1u, n / div);
            a.in_rows = in_rows_p();
            a.ptiles = ptiles.p;
            a.prow = prow.p;
            a.tpart = tpart.p;
            a.amask = amask.p;
            a.mask_max = std::max<uint32_t>(1u, n / kPullMaskDivisor);
            a.odeg = odeg.p;
        }
        return a;
    }

    // Pull plan capacity: every incoming row of kPullSplit slots or more is
    // cut into tiles of kPullTile slots, and rows hold at most the pools' slots.
    void ensure_pull() {
        const uint64_t slots = hbm_slots + host_slots;
        const uint64_t rows_long = std::min<uint64_t>(n, slots / kPullSplit);
        const uint64_t need = slots / kPullTile + rows_long + 1;
        // sized by n: a later cig_load_csr may bring a larger graph
        if (amask.count < (uint64_t(n) + 63) / 64) amask.alloc((uint64_t(n) + 63) / 64);
        if (odeg.count < n) odeg.alloc(n);
        if (need <= ptile_cap) return;
        ptiles.alloc(need);
        tpart.alloc(need);
        prow.alloc(std::min<uint64_t>(n, slots / kPullTile) + 1);
        ptile_cap = need;
    }

    // Runtime::Converge (pr_residual.cuh:171-187): at most pr_max_rounds
    // consume/scatter rounds; residual left over stays pending.
    // plan_ready: the batch amendment already built the pull plan of the
    // current incoming rows.
    void pr_converge(bool plan_ready = false) {
        if (pull_rows() && !coop) ensure_pull();
        PrArgs a = pr_args(cfg.pr_max_rounds);
        if (coop) {
            auto kernel = k_pr_loop<1>;
            static thread_local uint32_t grid = 0;
            if (!grid) grid = coop_grid_for(kernel);
            grid_coop = grid;
            launch_coop(kernel, a, grid);
        } else if (a.pull_min == kNone) {
            advance_round(run_chain(
                [&](uint64_t tag, uint32_t seq) {
                    k_pr_consume<<<grid_round, kBlock, 0, stream>>>(a, prog_dev, tag, seq);
                    timed_round([&] { k_pr_scatter<1><<<grid_round, kBlock, 0, stream>>>(a, seq); });
                },
                cfg.pr_max_rounds, "PageRank"));
        } else {
            // a dense round gathers along the incoming rows (k_pr_pull +
            // k_pr_pull_long), a sparse one scatters; each kernel checks
            // the round's frontier size and returns when it is not its form
            if (!plan_ready) pull_plan(a, nullptr);
            k_pr_degrees<<<grid_small, 256, 0, stream>>>(rows.p, n, odeg.p);
            advance_round(run_chain(
                [&](uint64_t tag, uint32_t seq) {
                    k_pr_consume<<<grid_pull, kBlock, 0, stream>>>(a, prog_dev, tag, seq);
                    k_pr_scatter<1><<<grid_round, kBlock, 0, stream>>>(a, seq);
                    timed_round([&] {
                        switch (qpl) {
                        case 1: k_pr_pull<1><<<grid_pull, kBlock, 0, stream>>>(a, seq); break;
                        case 4: k_pr_pull<4><<<grid_pull, kBlock, 0, stream>>>(a, seq); break;
                        default: k_pr_pull<2><<<grid_pull, kBlock, 0, stream>>>(a, seq); break;
                        }
                    });
                    k_pr_pull_long<<<grid_small, kBlock, 0, stream>>>(a, seq);
                },
                cfg.pr_max_rounds, "PageRank"));
        }
        HIPCHK(hipGetLastError());
    }

    void pr_residual_norm() {
        k_pr_l1<<<grid_small, 256, 0, stream>>>(res.p, n, stats.p);
        HIPCHK(hipGetLastError());
    }

    // ------------------------------------------------------------ init
    void init(const uint32_t *sources, uint32_t nsrc) {
        if (!loaded) throw HipError(CIG_ESTATE, "cig_load_csr must come first");
        flush_finish();  // the last batch's clear needs its counters, which the reset zeroes
        reset_ctl(false);
        if (is_pr()) {
            k_init_pr<<<grid_small, 256, 0, stream>>>(rank.p, comp.p, res.p, stamp.p, q1.p, n, ctl.p);
        } else if (cfg.algo == CIG_ALGO_CC) {
            k_init_traversal<<<grid_small, 256, 0, stream>>>(state.p, stamp.p, n, 1, 0, q1.p, ctl.p);
        } else {
            if (!sources || nsrc < 1 || nsrc > CIG_MAX_SOURCES)
                throw HipError(CIG_EINVAL, "SSSP/BFS take 1..64 sources");
            for (uint32_t i = 0; i < nsrc; ++i)
                if (sources[i] >= n) throw HipError(CIG_EINVAL, "source out of range");
            if (nsrc > 1) {
                msi_init(sources, nsrc);
                sync();
                initialized = true;
                return;
            }
            msi_release();
            source = sources[0];
            fill32(stamp.p, n, 0);
            k_init_traversal<<<grid_small, 256, 0, stream>>>(state.p, stamp.p, n, 0, source, q1.p, ctl.p);
        }
        HIPCHK(hipGetLastError());
        sync();
        initialized = true;
    }

    // ------------------------------------------------------------ stats
    // quick: an overlapped call whose side stream is ready (side_ready): the
    // start event comes first, nothing waits for the side stream and the
    // mutation counters are left to it.
    void begin_call(cig_batch_stats *st, bool quick = false) {
        if (st) std::memset(st, 0, sizeof(*st));
        if (quick) HIPCHK(hipEventRecord(ev[0], stream));
        else flush_finish();
        scan_timed = false;
        overlapped = false;
        round_timed = 0;
        round_launches = 0;
        inval_waves = 0;
        const uint32_t every = std::max<uint32_t>(1, cfg.kernel_timing_every);
        stages_now = calls % every == 0;
        timing_now = !rev_begin.empty() && stages_now;
        ++calls;
        reset_ctl(true, quick);
        if (!quick) HIPCHK(hipEventRecord(ev[0], stream));
    }

    // Time between two recorded events. A call that ends by polling its
    // published control block has not synchronised the stream: an event the
    // runtime has not seen complete yet is waited for first.
    static void event_ms(hipEvent_t a, hipEvent_t b, float *ms) {
        hipError_t r = hipEventElapsedTime(ms, a, b);
        if (r == hipErrorNotReady) {
            HIPCHK(hipEventSynchronize(b));
            r = hipEventElapsedTime(ms, a, b);
        }
        HIPCHK(r);
    }

    float elapsed(int a, int b) {
        float ms = 0;
        event_ms(ev[a], ev[b], &ms);
        return ms;
    }

    void reduce_stats() {
        k_reduce_stats<<<1, kWave, 0, stream>>>(stats.p, ctl.p);
        HIPCHK(hipGetLastError());
    }

    // The control block at the end of a call: published by the stream's last
    // kernel into mapped host memory and read by polling its sequence word
    // (a copy plus a stream synchronisation costs tens of microseconds of
    // wake-up between consecutive calls). When it returns, all of the call's
    // work on the engine stream is done.
    void read_ctl_published() {
        const uint64_t seq = ++pub_seq;
        k_publish_ctl<<<1, 64, 0, stream>>>(ctl.p, pub_dev, seq);
        HIPCHK(hipGetLastError());
        for (uint32_t spin = 0;; ++spin) {
            if (__atomic_load_n(&pub.p->seq, __ATOMIC_ACQUIRE) == seq) break;
            if ((spin & 1023) == 1023) {
                const hipError_t q = hipStreamQuery(stream);
                if (q != hipSuccess && q != hipErrorNotReady)
                    throw HipError(CIG_EDEVICE, std::string("end of call: ") + hipGetErrorString(q));
                if (q == hipSuccess && __atomic_load_n(&pub.p->seq, __ATOMIC_ACQUIRE) != seq)
                    throw HipError(CIG_EINTERNAL, "end of call: control block not published");
            }
            __builtin_ia32_pause();
        }
        std::memcpy(hctl.p, &pub.p->c, sizeof(Ctl));
    }

    void end_call(cig_batch_stats *st, int marks) {
        HIPCHK(hipEventRecord(ev[7], stream));
        reduce_stats();
        read_ctl_published();
        if (hctl.p->pr_invalid) throw HipError(CIG_EINTERNAL, "PR non-finite rank/residual");
        if (!is_pr() && !cfg.max_rounds) {
            const uint32_t r = hctl.p->round;
            if (hctl.p->qcount[r % 3] || hctl.p->tcount[r % 3])
                throw HipError(CIG_EINTERNAL, "fixed point did not converge within the round bound");
        }
        if (!st) return;
        const Ctl &c = *hctl.p;
        st->total_ms = elapsed(0, 7);
        if (marks >= 4 && stages_now) {  // ev: 0 start, 1 mutated, 2 invalidated, 3 repaired, 5 seeded, 4 propagated
            st->mutate_ms = elapsed(0, 1);
            // overlapped: the invalidation ran beside the mutation, from the start
            st->invalidate_ms = overlapped ? elapsed(0, 2) : elapsed(1, 2);
            st->repair_ms = overlapped ? elapsed(1, 3) : elapsed(2, 3);
            st->seed_ms = elapsed(3, 5);
            st->propagate_ms = elapsed(5, 4);
        }
        if (scan_timed) st->scan_kernel_ms = elapsed(8, 9);
        for (uint32_t i = 0; i < round_timed; ++i) {
            float ms = 0;
            event_ms(rev_begin[i], rev_end[i], &ms);
            st->round_kernel_ms += ms;
        }
        st->round_launches = round_launches;
        st->inval_rounds = inval_waves;
        for (int i = 0; i < 4; ++i) {
            st->stage_edges[i] = c.st_edges[i];
            st->stage_vertices[i] = c.st_vexp[i];
            st->stage_host_edges[i] = c.st_host[i];
        }
        st->edges_scanned = c.edges_scanned;
        st->host_edges_scanned = c.host_edges;
        st->vertices_expanded = c.vexp;
        st->affected = c.acount;
        st->rounds = c.rounds;
        st->deletes_applied = c.deletes_applied;
        st->missing_deletes = c.missing;
        st->invalid_updates = c.invalid;
        st->adds_applied = c.adds_applied;
        st->relocated_rows = c.relocated;
        st->touched_sources = c.touched_sources;
        st->converged = is_pr() ? (c.qcount[c.round % 3] == 0 ? 1u : 0u) : 1u;
        st->pr_residual_l1 = c.residual_l1;
        st->pull_rounds = c.pull_rounds;
    }

    // ------------------------------------------------------------ mutation
    MutArgs mut_args(const cig_edge *d, uint32_t nd, const cig_edge *a, uint32_t na) {
        MutArgs m;
        std::memset(&m, 0, sizeof(m));  // padding too: graph signatures compare the bytes
        m.ctl = ctl.p;
        m.rows = rows.p;
        m.pools = pools();
        m.fl = free_list();
        m.n = n;
        m.rev = rev ? 1u : 0u;
        m.hot = (cfg.flags & CIG_CFG_NO_HOTNESS) || !budget_slots ? nullptr : hot.p;
        m.dels = d;
        m.nd = nd;
        m.adds = a;
        m.na = na;
        m.cnt64 = cnt64.p;
        m.rslot = rslot.p;
        m.gofs = gofs.p;
        m.touched = touched.p;
        m.grouped = grouped.p;
        m.small_list = small_list.p;
        m.mid_list = mid_list.p;
        m.mid16_list = mid16_list.p;
        m.big_list = big_list.p;
        m.add_list = add_list.p;
        m.overflow = overflow.p;
        m.pieces = pieces.p;
        m.stats = stats.p;
        m.htoff = htoff.p;
        m.hkeys = hkeys.p;
        m.hcnts = hcnts.p;
        m.hcnt = hcnt.p;
        m.dclass = dclass.p;
        m.skip = cur_skip;
        m.dtiles = dtiles.p;
        return m;
    }

    // Overlapped batches: the deletes' claims run on the second stream beside
    // the invalidation waves too (n_stream >= 3, the default), or only the
    // grouping does (n_stream = 2).
    bool claims_beside() const { return cfg.n_stream == 0 || cfg.n_stream >= 3; }

    // A large batch (by its delete records, mirrors included) sends the HBM
    // medium rows with 9-16 deletes to the 16-request pass; graph-captured
    // calls, whose counts the host does not know, keep the plain form. The
    // grouping and the claims of one call take the same form.
    static bool mid16(const MutArgs &m) { return !m.ref && (uint64_t(m.nd) << m.rev) >= kMid16MinRecords; }

    void mutate_group(const MutArgs &m, hipStream_t on = nullptr) {
        if (!on) on = stream;
        // Grids sized by the call's records (all are grid-stride loops): a
        // grid of idle workgroups would hold the CUs that the invalidation
        // waves, running beside the grouping, need. Graph-captured calls read
        // their record counts on the device and keep the full grid.
        const uint64_t recs = (uint64_t(m.nd) + m.na) << m.rev;
        k_count<<<m.ref ? grid_small : grid_for(recs, 4), 256, 0, on>>>(m);
        if (mid16(m)) k_plan<true><<<m.ref ? grid_small : grid_for(recs, 1), 256, 0, on>>>(m);
        else k_plan<false><<<m.ref ? grid_small : grid_for(recs, 1), 256, 0, on>>>(m);
        k_scatter<<<m.ref ? grid_small : grid_for(recs, 1), 256, 0, on>>>(m);
        HIPCHK(hipGetLastError());
    }

    // The delete claims: each delete record's slot is found and set to
    // kNone (a tombstone) and listed as a hole; no live slot moves.
    void mutate_claims(const MutArgs &m, hipStream_t on = nullptr) {
        if (!on) on = stream;
        if (!m.nd) return;
        // tiny rows, host-tier medium rows (and in a large batch the HBM
        // medium rows with 9-16 deletes) in one launch, whose register needs
        // are alike; the HBM medium rows with up to 8 deletes keep their own
        if (mid16(m))
            k_del_lane_host<true><<<grid_small + grid_small / 2, kBlock, 0, on>>>(m, grid_small / 2, grid_small);
        else
            k_del_lane_host<false><<<grid_small, kBlock, 0, on>>>(m, grid_small / 2, grid_small);
        k_del_rows<4, false><<<grid_small, kBlock, 0, on>>>(m);
        k_del_scan<<<grid_small, 256, 0, on>>>(m);
        HIPCHK(hipGetLastError());
    }

    // claimed: mutate_claims ran already (the second stream of an overlapped
    // batch); then only the hole fixes and the adds are left.
    void mutate_apply(const MutArgs &m, bool claimed = false) {
        if (m.nd) {
            if (!claimed) mutate_claims(m);
            k_del_fix_both<<<grid_small, 256, 0, stream>>>(m, grid_small / 2);
        }
        k_add_plan<<<grid_small, 256, 0, stream>>>(m, add_list.p, &ctl.p->add_cnt);
        k_add_copy<<<grid_small, kBlock, 0, stream>>>(m);
        if (m.fl.off) k_fl_release<<<grid_small, 256, 0, stream>>>(m);
        HIPCHK(hipGetLastError());
    }

    void mutate_finish(const MutArgs &m, bool clear_affected = false) {
        k_clear_touched<<<grid_small, 256, 0, stream>>>(m, clear_affected ? aff.p : nullptr, alist.p);
        HIPCHK(hipGetLastError());
    }

    // A traversal batch ends with the host's read of the control block; its
    // end-of-call clear (per-row counters, affected bits, activity counts) is
    // enqueued at the start of the next call instead, ahead of that call's
    // control reset, so that the stream does not idle through a host round
    // trip. Everything that needs the counters zero begins with flush_finish.
    bool finish_pending = false;
    MutArgs pending_m{};
    bool pending_aff = false;

    // End of an overlapped call: the per-row counts and the mutation counters
    // are cleared on the side stream (which groups the next call's records),
    // the affected bits on the engine stream (which runs the next waves).
    // Called after the call's final synchronisation, so the engine stream is
    // idle. The clear also adds the activity counts (hot), which the next
    // call's rounds add to as well: with activity on, the next call is not
    // quick and waits for it (flush_finish).
    void finish_split(const MutArgs &m) {
        k_clear_touched<<<grid_small, 256, 0, side>>>(m, nullptr, nullptr);
        k_reset_mut<<<1, 64, 0, side>>>(ctl.p);
        HIPCHK(hipEventRecord(side_done_ev, side));
        k_clear_aff<<<grid_small, 256, 0, stream>>>(ctl.p, aff.p, alist.p);
        HIPCHK(hipGetLastError());
        side_pending = true;
        side_ready = m.hot == nullptr;
    }

    void defer_finish(const MutArgs &m, bool clear_affected) {
        pending_m = m;
        pending_aff = clear_affected;
        finish_pending = true;
    }

    void flush_finish() {
        if (side_pending) {
            HIPCHK(hipStreamWaitEvent(stream, side_done_ev, 0));
            side_pending = false;
        }
        side_ready = false;
        if (!finish_pending) return;
        finish_pending = false;
        if (msk) msi_clear();
        mutate_finish(pending_m, pending_aff);
    }

    // Rows that found no room: reclaim garbage (compaction), demote to the
    // host tier if HBM is still short, then append their adds again.
    DBuf<RowCount> ovf_rows;  // descriptors and counts of the rows whose adds found no room
    bool recover_overflow(const MutArgs &m) {
        read_ctl();
        const uint32_t nover = hctl.p->overflow_cnt;
        if (!nover) return false;
        std::vector<uint32_t> list(nover);
        HIPCHK(hipMemcpyAsync(list.data(), overflow.p, nover * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        // only the overflowing rows' descriptors and counts cross to the host
        // (a copy of every descriptor is 4 GB at scale 27)
        grow(ovf_rows, nover);
        k_gather_rows<<<grid_for(nover, 1), 256, 0, stream>>>(rows.p, cnt64.p, overflow.p, nover, ovf_rows.p);
        HIPCHK(hipGetLastError());
        std::vector<RowCount> rc(nover);
        HIPCHK(hipMemcpyAsync(rc.data(), ovf_rows.p, nover * sizeof(RowCount), hipMemcpyDeviceToHost, stream));
        sync();
        uint64_t need[2] = {0, 0};
        for (uint32_t i = 0; i < nover; ++i) {
            const Row &r = rc[i].row;
            need[(r.off & kHostTier) ? 1 : 0] += next_pow2_cap(uint64_t(r.deg) + uint32_t(rc[i].cnt >> 32));
        }
        for (int tier = 0; tier < 2; ++tier)
            if (need[tier]) compact_pool(tier, need[tier]);
        // reset the overflow list and copy jobs; retry the adds of those rows
        HIPCHK(hipMemcpyAsync(add_list.p, list.data(), nover * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        uint32_t zero = 0;
        HIPCHK(hipMemcpyAsync(&ctl.p->add_cnt, &nover, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(&ctl.p->overflow_cnt, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(&ctl.p->copy_cnt, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        MutArgs mm = mut_args(m.dels, m.nd, m.adds, m.na);  // pools/rows may have moved
        ensure_pieces(uint64_t(m.nd) << m.rev, uint64_t(m.na) << m.rev);
        mm.pieces = pieces.p;
        k_add_plan<<<grid_small, 256, 0, stream>>>(mm, add_list.p, &ctl.p->add_cnt);
        k_add_copy<<<grid_small, kBlock, 0, stream>>>(mm);
        if (mm.fl.off) k_fl_release<<<grid_small, 256, 0, stream>>>(mm);
        HIPCHK(hipGetLastError());
        read_ctl();
        if (hctl.p->overflow_cnt) throw HipError(CIG_ENOMEM, "edge pool exhausted after compaction");
        return true;
    }

    // Compacts one tier's pool (dropping relocation garbage) into a new
    // allocation with room for `extra` more slots. HBM is bounded by the
    // cache budget: rows that do not fit are demoted to the host tier.
    uint64_t host_compactions = 0;
    uint64_t refreshes_deferred = 0;  // refreshes that moved nothing rather than compact the host tier
    double compaction_ms[2] = {0.0, 0.0};

    void compact_pool(int tier, uint64_t extra) {
        const auto t0 = std::chrono::steady_clock::now();
        struct Clock {  // wall time of this compaction, whichever way it ends
            std::chrono::steady_clock::time_point t0;
            double *acc;
            ~Clock() {
                *acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
        } clock{t0, &compaction_ms[tier]};
        if (tier == 1) ++host_compactions;
        reduce_stats();  // garbage counted so far lands in the control block first
        if (compact_in_place(tier, extra)) {
            ++compactions;
            return;
        }
        sync();
        std::vector<Row> hrows(nrows);
        HIPCHK(hipMemcpy(hrows.data(), rows.p, size_t(nrows) * sizeof(Row), hipMemcpyDeviceToHost));
        auto in_tier = [&](const Row &r) { return ((r.off & kHostTier) != 0) == (tier == 1); };
        uint64_t live = 0;
        for (const Row &r : hrows)
            if (in_tier(r)) live += r.cap;
        if (tier == 0 && budget_slots && live + extra > budget_slots) {
            demote_for_space(hrows, live + extra - budget_slots);
            live = 0;
            for (const Row &r : hrows)
                if (in_tier(r)) live += r.cap;
        }
        uint64_t cap = live + extra + live / 4 + (uint64_t(1) << 20);
        if (tier == 0 && budget_slots) cap = std::max<uint64_t>(budget_slots, live + extra);
        std::vector<Row> moved = hrows;
        uint64_t tail = 0;
        for (uint32_t v = 0; v < nrows; ++v) {
            if (!in_tier(hrows[v])) continue;
            moved[v].off = tail | (tier ? kHostTier : 0ull);
            tail += hrows[v].cap;
        }
        if (tier == 0) {
            DBuf<uint32_t> np;
            np.alloc(cap);
            DBuf<Row> drows;
            drows.alloc(nrows);
            HIPCHK(hipMemcpy(drows.p, moved.data(), size_t(nrows) * sizeof(Row), hipMemcpyHostToDevice));
            k_relocate_rows<<<grid_small, 256, 0, stream>>>(rows.p, drows.p, nrows, 0, hbm_pool.p, np.p);
            HIPCHK(hipGetLastError());
            sync();
            std::swap(hbm_pool.p, np.p);
            std::swap(hbm_pool.count, np.count);
            std::swap(rows.p, drows.p);
            hbm_slots = cap;
        } else {
            HBuf<uint32_t> np;
            np.alloc(cap);
            parallel_for(nrows, pick_threads(0), [&](uint64_t lo, uint64_t hi, unsigned) {
                for (uint64_t v = lo; v < hi; ++v) {
                    if (!in_tier(hrows[v]) || !hrows[v].deg) continue;
                    std::memcpy(np.p + (moved[v].off & ~kHostTier), host_pool.p + (hrows[v].off & ~kHostTier),
                                size_t(hrows[v].deg) * sizeof(uint32_t));
                }
            });
            HIPCHK(hipMemcpy(rows.p, moved.data(), size_t(nrows) * sizeof(Row), hipMemcpyHostToDevice));
            std::swap(host_pool.p, np.p);
            std::swap(host_pool.count, np.count);
            host_pool_dev = host_pool.device_ptr();
            host_slots = cap;
            clear_free_list();
        }
        uint64_t vals[2] = {tail, cap};
        uint64_t zero = 0;
        HIPCHK(hipMemcpy(&ctl.p->pool_tail[tier], &vals[0], sizeof(uint64_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(&ctl.p->pool_cap[tier], &vals[1], sizeof(uint64_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(&ctl.p->garbage[tier], &zero, sizeof(uint64_t), hipMemcpyHostToDevice));
        ensure_tiles();
        ++compactions;
    }

    // In-place compaction of one tier on the device (kernels k_cmp_*): rows are
    // bucketed by 2^k-slot windows of their offset, each window packed through
    // a staging buffer. The host reads only the per-window totals (at most
    // kCmpMaxWin of them). Returns false, changing nothing, when the live rows
    // plus `extra` do not fit the tier's pool (or the HBM budget): the caller
    // then grows the pool or demotes rows.
    DBuf<unsigned long long> cmp_win;  // wcnt | wcap | lfill | cursor, kCmpMaxWin each
    DBuf<uint32_t> cmp_list, cmp_stage;
    DBuf<uint32_t> cmp_pc, cmp_pjob;  // per window: copy pieces; piece -> job
    DBuf<CmpJob> cmp_jobs;
    uint64_t inplace_compactions = 0;

    bool compact_in_place(int tier, uint64_t extra) {
        if (tier == 1) return compact_host_in_place(extra);
        read_ctl();
        const uint64_t tail = hctl.p->pool_tail[tier], cap = hctl.p->pool_cap[tier];
        if (tail == 0) return false;
        uint64_t wslots = uint64_t(1) << 22;
        while ((tail + wslots - 1) / wslots > kCmpMaxWin) wslots <<= 1;
        const uint32_t nwin = uint32_t((tail + wslots - 1) / wslots);
        if (!cmp_win.p) cmp_win.alloc(4 * kCmpMaxWin);
        if (!cmp_pc.p) cmp_pc.alloc(kCmpMaxWin);
        fill32(reinterpret_cast<uint32_t *>(cmp_win.p), 8ull * kCmpMaxWin, 0u);
        fill32(cmp_pc.p, kCmpMaxWin, 0u);
        CmpArgs a{};
        a.rows = rows.p;
        a.nrows = nrows;
        a.tier = uint32_t(tier);
        a.nwin = nwin;
        a.wslots = wslots;
        a.wcnt = cmp_win.p;
        a.wcap = cmp_win.p + kCmpMaxWin;
        a.lfill = cmp_win.p + 2 * kCmpMaxWin;
        a.pools = pools();
        const uint32_t chunks = (nrows + kCmpChunk - 1) / kCmpChunk;
        k_cmp_hist<<<chunks, 256, 0, stream>>>(a);
        HIPCHK(hipGetLastError());
        std::vector<unsigned long long> h(2 * kCmpMaxWin);
        HIPCHK(hipMemcpyAsync(h.data(), cmp_win.p, h.size() * 8, hipMemcpyDeviceToHost, stream));
        sync();
        uint64_t live = 0, rowsn = 0, maxcnt = 0, maxcap = 0, maxpieces = 0;
        std::vector<unsigned long long> lbase(kCmpMaxWin, 0), nbase(nwin, 0);
        for (uint32_t w = 0; w < nwin; ++w) {
            lbase[w] = rowsn;
            nbase[w] = live;
            rowsn += h[w];
            live += h[kCmpMaxWin + w];
            maxcnt = std::max<uint64_t>(maxcnt, h[w]);
            maxcap = std::max<uint64_t>(maxcap, h[kCmpMaxWin + w]);
            // a row has max(1, ceil(deg / kCmpPiece)) pieces, deg <= cap
            maxpieces = std::max<uint64_t>(maxpieces, h[w] + h[kCmpMaxWin + w] / kCmpPiece + 1);
        }
        const uint64_t limit = tier == 0 && budget_slots ? std::min(cap, budget_slots) : cap;
        if (live + extra > limit || maxcnt > 0xFFFFFFFFull || maxpieces > 0xFFFFFFFFull) return false;
        grow(cmp_list, rowsn);
        grow(cmp_jobs, maxcnt);
        grow(cmp_stage, maxcap);
        grow(cmp_pjob, maxpieces);
        a.list = cmp_list.p;
        HIPCHK(hipMemcpyAsync(a.lfill, lbase.data(), kCmpMaxWin * 8, hipMemcpyHostToDevice, stream));
        k_cmp_scatter<<<chunks, 256, 0, stream>>>(a);
        HIPCHK(hipGetLastError());
        unsigned long long *cursor = cmp_win.p + 3 * kCmpMaxWin;
        for (uint32_t w = 0; w < nwin; ++w) {
            const uint32_t cnt = uint32_t(h[w]);
            if (!cnt) continue;
            k_cmp_plan<<<grid_small, 256, 0, stream>>>(a, cmp_list.p + lbase[w], cnt, cursor + w, cmp_pc.p + w,
                                                       cmp_jobs.p, cmp_pjob.p);
            k_cmp_in<<<grid_small, 256, 0, stream>>>(a, cmp_jobs.p, cmp_pjob.p, cmp_pc.p + w, cmp_stage.p);
            k_cmp_out<<<grid_small, 256, 0, stream>>>(a, cmp_jobs.p, cmp_pjob.p, cmp_pc.p + w, cmp_stage.p,
                                                      nbase[w]);
            HIPCHK(hipGetLastError());
        }
        const uint64_t zero = 0;
        HIPCHK(hipMemcpyAsync(&ctl.p->pool_tail[tier], &live, sizeof(uint64_t), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(&ctl.p->garbage[tier], &zero, sizeof(uint64_t), hipMemcpyHostToDevice, stream));
        sync();
        ++inplace_compactions;
        return true;
    }

    // The host tier packed in place by host threads on the pinned pool, with
    // the same windows as compact_in_place. Copying host-tier rows through the
    // device would cross the host link twice per row, one small row per wave
    // step (latency-bound: minutes for the 10^8 short rows of a scale-27
    // graph); host threads move them at memory bandwidth instead. The row
    // descriptors make one round trip; nothing else is allocated but a
    // staging buffer of one window.
    bool compact_host_in_place(uint64_t extra) {
        reduce_stats();
        read_ctl();
        const uint64_t tail = hctl.p->pool_tail[1], cap = hctl.p->pool_cap[1];
        if (tail == 0) return false;
        uint64_t wslots = uint64_t(1) << 22;
        while ((tail + wslots - 1) / wslots > kCmpMaxWin) wslots <<= 1;
        const uint32_t nwin = uint32_t((tail + wslots - 1) / wslots);
        std::vector<Row> hrows(nrows);
        HIPCHK(hipMemcpy(hrows.data(), rows.p, size_t(nrows) * sizeof(Row), hipMemcpyDeviceToHost));
        const unsigned T = pick_threads(0);
        auto in_host = [](const Row &r) { return r.cap != 0 && (r.off & kHostTier); };
        // per-thread window histograms, then row ids grouped by window
        std::vector<uint64_t> tcnt(size_t(T) * nwin, 0), tcap(size_t(T) * nwin, 0);
        parallel_for(nrows, T, [&](uint64_t lo, uint64_t hi, unsigned t) {
            for (uint64_t v = lo; v < hi; ++v)
                if (in_host(hrows[v])) {
                    const uint64_t w = (hrows[v].off & ~kHostTier) / wslots;
                    tcnt[size_t(t) * nwin + w]++;
                    tcap[size_t(t) * nwin + w] += hrows[v].cap;
                }
        });
        std::vector<uint64_t> wbase(nwin + 1, 0), nbase(nwin + 1, 0), wcap(nwin, 0);
        uint64_t rowsn = 0, live = 0, maxcap = 0;
        for (uint32_t w = 0; w < nwin; ++w) {
            wbase[w] = rowsn;
            nbase[w] = live;
            for (unsigned t = 0; t < T; ++t) {
                const uint64_t c = tcnt[size_t(t) * nwin + w];
                tcnt[size_t(t) * nwin + w] = rowsn;  // this thread's first slot in the window's list
                rowsn += c;
                wcap[w] += tcap[size_t(t) * nwin + w];
            }
            live += wcap[w];
            maxcap = std::max(maxcap, wcap[w]);
        }
        wbase[nwin] = rowsn;
        if (live + extra > cap) return false;
        std::vector<uint32_t> list(rowsn);
        parallel_for(nrows, T, [&](uint64_t lo, uint64_t hi, unsigned t) {
            for (uint64_t v = lo; v < hi; ++v)
                if (in_host(hrows[v])) list[tcnt[size_t(t) * nwin + (hrows[v].off & ~kHostTier) / wslots]++] = uint32_t(v);
        });
        // One pool of T threads walks the windows in order; per window each
        // thread takes a contiguous slice of the window's rows: it sums their
        // capacities (the slices' sums give each slice its base), copies its
        // rows into the staging buffer, and after a barrier (every row of the
        // window read) writes them to their packed places. A window whose
        // packed range ends at or below its own first slot (the garbage before
        // it is at least its capacity) overlaps no unread row: its rows are
        // copied straight to their places, with no staging pass.
        std::vector<uint32_t> stage(maxcap);
        std::vector<uint64_t> part(size_t(T) + 1, 0);
        SpinBarrier team_bar(T), solo_bar(1);
        uint32_t *pool = host_pool.p;
        run_team(T, [&](unsigned t, unsigned team) {
            SpinBarrier &bar = team == T ? team_bar : solo_bar;
            for (uint32_t w = 0; w < nwin; ++w) {
                const uint64_t b = wbase[w], e = wbase[w + 1];
                if (b == e) continue;
                const uint64_t chunk = (e - b + team - 1) / team;
                const uint64_t lo = std::min(e, b + t * chunk), hi = std::min(e, lo + chunk);
                uint64_t sum = 0;
                for (uint64_t i = lo; i < hi; ++i) sum += hrows[list[i]].cap;
                part[t + 1] = sum;
                bar.wait();
                uint64_t at = 0;
                for (unsigned j = 1; j <= t; ++j) at += part[j];
                const uint64_t first = at;
                if (nbase[w] + wcap[w] <= uint64_t(w) * wslots) {
                    for (uint64_t i = lo; i < hi; ++i) {
                        Row &r = hrows[list[i]];
                        const uint64_t from = r.off & ~kHostTier;
                        r.off = (nbase[w] + at) | kHostTier;
                        std::memcpy(pool + (nbase[w] + at), pool + from, size_t(r.deg) * 4);
                        at += r.cap;
                    }
                    bar.wait();  // part is free for the next window
                    continue;
                }
                for (uint64_t i = lo; i < hi; ++i) {
                    const Row &r = hrows[list[i]];
                    std::memcpy(stage.data() + at, pool + (r.off & ~kHostTier), size_t(r.deg) * 4);
                    at += r.cap;
                }
                bar.wait();  // the window's rows are all in the staging buffer
                at = first;
                for (uint64_t i = lo; i < hi; ++i) {
                    Row &r = hrows[list[i]];
                    r.off = (nbase[w] + at) | kHostTier;
                    std::memcpy(pool + (nbase[w] + at), stage.data() + at, size_t(r.deg) * 4);
                    at += r.cap;
                }
                bar.wait();  // part and the staging buffer are free for the next window
            }
        });
        HIPCHK(hipMemcpy(rows.p, hrows.data(), size_t(nrows) * sizeof(Row), hipMemcpyHostToDevice));
        const uint64_t zero = 0;
        HIPCHK(hipMemcpy(&ctl.p->pool_tail[1], &live, sizeof(uint64_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(&ctl.p->garbage[1], &zero, sizeof(uint64_t), hipMemcpyHostToDevice));
        clear_free_list();  // the listed blocks lay in the garbage just packed away
        ++inplace_compactions;
        return true;
    }

    // Moves the coldest HBM rows (lowest activity score, then shortest) to the
    // host tier until `slots` are freed; updates hrows and the device rows.
    void demote_for_space(std::vector<Row> &hrows, uint64_t slots) {
        std::vector<uint32_t> sc(nrows, 0u);
        HIPCHK(hipMemcpy(sc.data(), score.p, size_t(nrows) * sizeof(uint32_t), hipMemcpyDeviceToHost));
        std::vector<uint32_t> cand;
        for (uint32_t v = 0; v < nrows; ++v)
            if (!(hrows[v].off & kHostTier) && hrows[v].cap) cand.push_back(v);
        std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) {
            return sc[a] != sc[b] ? sc[a] < sc[b] : hrows[a].deg < hrows[b].deg;
        });
        std::vector<uint32_t> move;
        uint64_t freed = 0, host_need = 0;
        for (uint32_t v : cand) {
            if (freed >= slots) break;
            move.push_back(v);
            freed += hrows[v].cap;
            host_need += hrows[v].cap;
        }
        move_rows(hrows, move, 1, host_need);
    }

    // Re-homes `vs` into `tier`, growing that tier's pool if needed.
    void move_rows(std::vector<Row> &hrows, const std::vector<uint32_t> &vs, int tier, uint64_t need) {
        if (vs.empty()) return;
        read_ctl();
        uint64_t tail = hctl.p->pool_tail[tier];
        uint64_t cap = hctl.p->pool_cap[tier];
        if (tail + need > cap) {
            compact_pool(tier, need);
            HIPCHK(hipMemcpy(hrows.data(), rows.p, size_t(nrows) * sizeof(Row), hipMemcpyDeviceToHost));
            read_ctl();
            tail = hctl.p->pool_tail[tier];
            cap = hctl.p->pool_cap[tier];
            if (tail + need > cap) throw HipError(CIG_ENOMEM, "tier pool cannot hold moved rows");
        }
        std::vector<CopyJob> hj;
        std::vector<uint32_t> hp;  // copy piece -> job
        for (uint32_t v : vs) {
            Row &r = hrows[v];
            CopyJob j{};
            j.from = r.off;
            j.to = tail | (tier ? kHostTier : 0ull);
            j.len = r.deg;
            j.pfirst = uint32_t(hp.size());
            const uint32_t np = std::max(1u, (r.deg + kCmpPiece - 1) / kCmpPiece);
            hp.insert(hp.end(), np, uint32_t(hj.size()));
            hj.push_back(j);
            r.off = j.to;
            tail += r.cap;
        }
        DBuf<CopyJob> dj;
        dj.alloc(hj.size());
        DBuf<uint32_t> dpj, dcount;
        dpj.alloc(hp.size());
        dcount.alloc(1);
        const uint32_t np = uint32_t(hp.size());
        HIPCHK(hipMemcpy(dj.p, hj.data(), hj.size() * sizeof(CopyJob), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dpj.p, hp.data(), hp.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dcount.p, &np, sizeof(uint32_t), hipMemcpyHostToDevice));
        k_move_rows<<<grid_small, 256, 0, stream>>>(dj.p, dpj.p, dcount.p, pools());
        HIPCHK(hipGetLastError());
        sync();
        HIPCHK(hipMemcpy(rows.p, hrows.data(), size_t(nrows) * sizeof(Row), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(&ctl.p->pool_tail[tier], &tail, sizeof(uint64_t), hipMemcpyHostToDevice));
        uint64_t g = 0;
        for (uint32_t v : vs) g += hrows[v].cap;
        // the vacated blocks of the other tier become garbage
        uint64_t og = 0;
        HIPCHK(hipMemcpy(&og, &ctl.p->garbage[1 - tier], sizeof(uint64_t), hipMemcpyDeviceToHost));
        og += g;
        HIPCHK(hipMemcpy(&ctl.p->garbage[1 - tier], &og, sizeof(uint64_t), hipMemcpyHostToDevice));
    }

    uint64_t compactions = 0;

    // ------------------------------------------------------------ traversal batch
    InvalArgs inval_args(const cig_edge *d, uint32_t nd, const cig_edge *a, uint32_t na) {
        InvalArgs x;
        std::memset(&x, 0, sizeof(x));
        x.ctl = ctl.p;
        x.rows = rows.p;
        x.in_rows = in_rows_p();
        x.state = state.p;
        x.stamp = stamp.p;
        x.aff = aff.p;
        x.alist = alist.p;
        x.filt = filt.p;
        x.q0 = q0.p;
        x.q1 = q1.p;
        x.t0 = t0.p;
        x.t1 = t1.p;
        x.cfg = ecfg(kStageInvalidate);
        x.cfg.adapt = adapt_split;
        x.dels = d;
        x.nd = nd;
        x.adds = a;
        x.na = na;
        x.n = n;
        x.kind = cfg.algo == CIG_ALGO_CC ? 1u : 0u;
        x.max_rounds = round_bound();  // each round consumes at least one new item
        x.seed = cfg.weight_seed;
        x.add_skip = cur_skip;
        x.del_skip = cur_skip ? cur_skip + na : nullptr;
        return x;
    }

    template <int W, bool S, int QPL>
    void repair_scan_qpl(const InvalArgs &ia) {
        InvalArgs a = ia;
        a.cfg = ecfg(kStageRepair);
        fill32(filt.p, kFilterWords, 0);
        k_build_filter<<<grid_small, 256, 0, stream>>>(a);
        // 32 KB filter + 8 wave slots of LDS per workgroup: 3 workgroups per CU
        const uint32_t grid = uint32_t(cus) * 3;
        mark(8);
        k_repair_scan<W, S, QPL><<<grid, kRepairBlock, 0, stream>>>(a);
        mark(9);
        scan_timed = stages_now;
        k_repair_tiles<W, S, QPL><<<grid, kRepairBlock, 0, stream>>>(a);
        HIPCHK(hipGetLastError());
    }

    template <int W, bool S>
    void repair_scan(const InvalArgs &ia) {
        switch (qpl) {
        case 1: repair_scan_qpl<W, S, 1>(ia); break;
        case 4: repair_scan_qpl<W, S, 4>(ia); break;
        default: repair_scan_qpl<W, S, 2>(ia); break;
        }
    }

    template <int W, bool S, int QPL>
    void repair_pull_qpl(const InvalArgs &ia) {
        InvalArgs a = ia;
        a.cfg = ecfg(kStageRepair);
        mark(8);
        k_repair_pull<W, S, QPL><<<grid_small, kBlock, 0, stream>>>(a);
        mark(9);
        scan_timed = stages_now;
        k_repair_pull_tiles<W, S, QPL><<<grid_small, kBlock, 0, stream>>>(a);
        HIPCHK(hipGetLastError());
    }

    // Deletion repair: pull over the incoming rows of the affected vertices,
    // or (without the incoming index) a push scan over every row.
    template <int W, bool S>
    void repair(const InvalArgs &ia) {
        if (!pull_rows()) return repair_scan<W, S>(ia);
        switch (qpl) {
        case 1: repair_pull_qpl<W, S, 1>(ia); break;
        case 4: repair_pull_qpl<W, S, 4>(ia); break;
        default: repair_pull_qpl<W, S, 2>(ia); break;
        }
    }

    template <int W, bool S>
    void insert_seed(const InvalArgs &ia, bool skip_affected) {
        k_insert_seed<W, S><<<grid_small, 256, 0, stream>>>(ia, skip_affected ? 1u : 0u);
        HIPCHK(hipGetLastError());
    }

    void invalidate_and_reset(const InvalArgs &ia, hipStream_t on = nullptr, bool reset = true) {
        if (!on) on = stream;
        k_inval_seed<<<ia.ref ? grid_small : grid_for(ia.nd, 1), 256, 0, on>>>(ia);
        HIPCHK(hipGetLastError());
        InvalArgs a = ia;
        if (coop) {
            auto kernel = k_inval_loop<1>;
            static thread_local uint32_t grid = 0;
            if (!grid) grid = coop_grid_for(kernel);
            launch_coop(kernel, a, grid);
        } else {
            inval_waves += run_chain(
                [&](uint64_t tag, uint32_t seq) {
                    k_inval_round<1><<<grid_round, kBlock, 0, on>>>(a, prog_dev, tag, seq);
                },
                round_bound(), "invalidation", on);
        }
        if (reset) k_reset_affected<<<grid_small, 256, 0, on>>>(ia);
        HIPCHK(hipGetLastError());
    }

    // ------------------------------------------------ graph-replayed batches
    // CIG_CFG_GRAPH_BATCHES: the fixed-length parts of a batch run as two
    // captured HIP graphs, replayed for every batch: (1) the mutation and the
    // invalidation seeds, (2) the reset of the affected vertices, the pull
    // repair and the insertion seeds. Their kernels read the batch's records
    // through a BatchRef that one small kernel fills before each replay; the
    // round chains between and after them stay host-driven (their length
    // depends on the data). A graph is captured again when any buffer or
    // launch shape its arguments name has changed (the signature below).
    struct Captured {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        std::vector<uint8_t> sig;
        void reset() {
            if (x) (void)hipGraphExecDestroy(x);
            if (g) (void)hipGraphDestroy(g);
            x = nullptr;
            g = nullptr;
            sig.clear();
        }
        ~Captured() { reset(); }
    };
    Captured graph_mut, graph_repair;
    Captured graph_group, graph_apply;  // the overlapped form: side-stream grouping, engine-stream rest
    DBuf<BatchRef> bref;
    uint64_t graph_captures = 0, graph_replays = 0;

    bool graph_mode(bool staged) const {
        return (cfg.flags & CIG_CFG_GRAPH_BATCHES) && !coop && !msk && !uf_mode() && !is_pr() && !staged &&
               !cur_skip && pull_rows();
    }

    template <class T>
    static void sig_add(std::vector<uint8_t> &sig, const T &x) {
        const uint8_t *b = reinterpret_cast<const uint8_t *>(&x);
        sig.insert(sig.end(), b, b + sizeof(T));
    }

    template <class F>
    void graph_run(Captured &c, const std::vector<uint8_t> &sig, F body, hipStream_t on = nullptr) {
        if (!on) on = stream;
        if (!c.x || c.sig != sig) {
            c.reset();
            HIPCHK(hipStreamBeginCapture(on, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(on, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(on, &c.g));
            HIPCHK(hipGraphInstantiate(&c.x, c.g, nullptr, nullptr, 0));
            c.sig = sig;
            ++graph_captures;
        }
        HIPCHK(hipGraphLaunch(c.x, on));
        ++graph_replays;
    }

    // The engine-stream graph of an overlapped batch, launched once both the
    // invalidation chain and the side stream's grouping are done: the reset
    // of the affected vertices, the deletes and adds, the pull repair and
    // the insertion seeds.
    template <int W, bool S, int QPL>
    void graph_apply_body(const MutArgs &mc, const InvalArgs &ic) {
        k_reset_affected<<<grid_small, 256, 0, stream>>>(ic);
        HIPCHK(hipGetLastError());
        mutate_apply(mc, claims_beside());
        graph_repair_body<W, S, QPL>(ic, false);
    }

    template <int W, bool S, int QPL>
    void graph_repair_body(const InvalArgs &ia, bool reset = true) {
        InvalArgs a = ia;
        a.cfg = ecfg(kStageRepair);
        if (reset) k_reset_affected<<<grid_small, 256, 0, stream>>>(ia);
        k_repair_pull<W, S, QPL><<<grid_small, kBlock, 0, stream>>>(a);
        k_repair_pull_tiles<W, S, QPL><<<grid_small, kBlock, 0, stream>>>(a);
        k_insert_seed<W, S><<<grid_small, 256, 0, stream>>>(ia, 1u);  // no vertex is affected without deletes
        HIPCHK(hipGetLastError());
    }

    template <int W, bool S>
    void traversal_batch_graph(const cig_edge *d_dels, uint32_t nd, const cig_edge *d_adds, uint32_t na,
                               cig_batch_stats *st) {
        // with deletes and a second stream, the grouping replays on the side
        // stream beside the invalidation chain (as traversal_batch runs it)
        const bool overlap = nd && cfg.n_stream != 1;
        const bool quick = overlap && side_ready && fresh_records;
        fresh_records = false;
        side_ready = false;
        begin_call(st, quick);
        if (!bref.p) bref.alloc(1);
        if (overlap && !quick) {
            HIPCHK(hipEventRecord(fork_ev, stream));
            HIPCHK(hipStreamWaitEvent(side, fork_ev, 0));
        }
        // the side stream's graph reads the batch reference first; the engine
        // stream's graph reads it after the join
        k_set_batch<<<1, 64, 0, overlap ? side : stream>>>(bref.p, d_dels, nd, d_adds, na);
        HIPCHK(hipGetLastError());
        MutArgs m = mut_args(d_dels, nd, d_adds, na);
        InvalArgs ia = inval_args(d_dels, nd, d_adds, na);
        // the captured launches: every kernel (nd, na > 0) reading the records through bref
        MutArgs mc = mut_args(nullptr, 1, nullptr, 1);
        mc.ref = bref.p;
        InvalArgs ic = inval_args(nullptr, 1, nullptr, 1);
        ic.ref = bref.p;
        std::vector<uint8_t> sig;
        sig_add(sig, mc);
        sig_add(sig, ic);
        sig_add(sig, grid_small);
        sig_add(sig, qpl);
        if (overlap) {
            overlapped = true;
            graph_run(graph_group, sig, [&] {
                mutate_group(mc, side);
                if (claims_beside()) mutate_claims(mc, side);
            }, side);
            HIPCHK(hipEventRecord(join_ev, side));
            try {
                k_inval_seed<<<grid_for(nd, 1), 256, 0, stream>>>(ia);
                HIPCHK(hipGetLastError());
                inval_waves += run_chain([&](uint64_t tag, uint32_t seq) {
                              k_inval_round<1><<<grid_round, kBlock, 0, stream>>>(ia, prog_dev, tag, seq);
                          },
                          round_bound(), "invalidation");
            } catch (...) {
                (void)hipStreamSynchronize(side);
                throw;
            }
            mark(2);
            HIPCHK(hipStreamWaitEvent(stream, join_ev, 0));
            graph_run(graph_apply, sig, [&] {
                switch (qpl) {
                case 1: graph_apply_body<W, S, 1>(mc, ic); break;
                case 4: graph_apply_body<W, S, 4>(mc, ic); break;
                default: graph_apply_body<W, S, 2>(mc, ic); break;
                }
            });
            // deletes, adds and the repair replay as one graph: their stage
            // times are reported together as the mutation's
            mark(1);
            mark(3);
            mark(5);
            relax();
            mark(4);
            end_call(st, 4);
            if (na && hctl.p->overflow_cnt) {
                recover_overflow(m);
                ia.ref = bref.p;
                insert_seed<W, S>(ia, false);
                relax();
                end_call(st, 4);
            }
            finish_split(m);
            return;
        }
        graph_run(graph_mut, sig, [&] {
            mutate_group(mc);
            mutate_apply(mc);
            k_inval_seed<<<grid_small, 256, 0, stream>>>(ic);
            HIPCHK(hipGetLastError());
        });
        mark(1);
        if (nd) {
            inval_waves += run_chain([&](uint64_t tag, uint32_t seq) {
                          k_inval_round<1><<<grid_round, kBlock, 0, stream>>>(ia, prog_dev, tag, seq);
                      },
                      round_bound(), "invalidation");
        }
        mark(2);
        graph_run(graph_repair, sig, [&] {
            switch (qpl) {
            case 1: graph_repair_body<W, S, 1>(ic); break;
            case 4: graph_repair_body<W, S, 4>(ic); break;
            default: graph_repair_body<W, S, 2>(ic); break;
            }
        });
        mark(3);
        mark(5);
        relax();
        mark(4);
        end_call(st, 4);
        if (na && hctl.p->overflow_cnt) {
            recover_overflow(m);
            ia.ref = bref.p;
            insert_seed<W, S>(ia, false);
            relax();
            end_call(st, 4);
        }
        defer_finish(m, nd != 0);
    }

    template <int W, bool S>
    void traversal_batch(const cig_edge *d_dels, uint32_t nd, const cig_edge *d_adds, uint32_t na,
                         bool staged, cig_batch_stats *st) {
        if (graph_mode(staged)) return traversal_batch_graph<W, S>(d_dels, nd, d_adds, na, st);
        const bool overlap = nd && !coop && !(staged && na) && cfg.n_stream != 1;
        // quick: the side stream's counters are reset and this call enqueued
        // nothing before its batch, so the grouping needs no fork
        const bool quick = overlap && side_ready && fresh_records;
        fresh_records = false;
        side_ready = false;  // set again only once this call has finished (finish_split)
        begin_call(st, quick);
        if (staged && nd && na) {
            // stage 1: deletes to their fixed point, then stage 2: adds
            delete_stage<W, S>(d_dels, nd);
            insert_stage<W, S>(d_adds, na);
            mark(4);
            end_call(st, 0);
            return;
        }
        MutArgs m = mut_args(d_dels, nd, d_adds, na);
        InvalArgs ia = inval_args(d_dels, nd, d_adds, na);
        if (overlap) {
            // The invalidation waves read the rows as they were before the
            // batch: a child y of x (parent(y) = x) is found over any edge
            // x->y, and one whose edges x->y are all deleted is a seed
            // anyway, so the affected set is the one the mutated rows give.
            // The records are grouped (k_count / k_plan / k_scatter, which
            // only read the rows) and the deletes claim their slots (a slot
            // becomes kNone; nothing moves) on the side stream while the
            // waves run on the engine stream: a wave reads each slot either
            // before its claim (the deleted edge x->y, whose y is a seed when
            // it is a tree edge) or after it (kNone, which the waves skip).
            // The hole fixes, which move slots, and the adds follow both. The
            // side work is enqueued first, so that its join event exists
            // before the host drives the waves' chain.
            overlapped = true;
            if (!quick) {
                HIPCHK(hipEventRecord(fork_ev, stream));
                HIPCHK(hipStreamWaitEvent(side, fork_ev, 0));
            }
            mutate_group(m, side);
            const bool side_claims = claims_beside();
            if (side_claims) mutate_claims(m, side);
            HIPCHK(hipEventRecord(join_ev, side));
            try {
                invalidate_and_reset(ia, nullptr, false);
            } catch (...) {
                // the grouping may still run: nothing later may touch its counters
                (void)hipStreamSynchronize(side);
                throw;
            }
            // the join first: the host's wait call then overlaps the reset
            HIPCHK(hipStreamWaitEvent(stream, join_ev, 0));
            k_reset_affected<<<grid_small, 256, 0, stream>>>(ia);
            HIPCHK(hipGetLastError());
            mark(2);
            mutate_apply(m, side_claims);
            mark(1);
        } else {
            mutate_group(m);
            mutate_apply(m);
            mark(1);
            if (nd) invalidate_and_reset(ia);
            mark(2);
        }
        if (nd) repair<W, S>(ia);
        mark(3);
        if (na) insert_seed<W, S>(ia, nd != 0);
        mark(5);
        relax();
        mark(4);
        // One host read per batch: end_call's control-block read also tells
        // whether any insert found no room (then recover, re-seed those
        // inserts and converge again, and refresh the statistics).
        end_call(st, 4);
        if (na && hctl.p->overflow_cnt) {
            recover_overflow(m);
            insert_seed<W, S>(ia, false);
            relax();
            end_call(st, 4);  // the control block still holds the whole call's totals
        }
        if (overlap) finish_split(m);
        else defer_finish(m, nd != 0);
    }

    template <int W, bool S>
    void delete_stage(const cig_edge *d_dels, uint32_t nd) {
        MutArgs m = mut_args(d_dels, nd, nullptr, 0);
        mutate_group(m);
        mutate_apply(m);
        InvalArgs ia = inval_args(d_dels, nd, nullptr, 0);
        invalidate_and_reset(ia);
        repair<W, S>(ia);
        relax();
        mutate_finish(m, true);
    }

    template <int W, bool S>
    void insert_stage(const cig_edge *d_adds, uint32_t na) {
        // counters of the previous stage stay in the control block; only the
        // per-mutation lists are cleared
        flush_finish();
        k_reset_mut<<<1, 64, 0, stream>>>(ctl.p);
        MutArgs m = mut_args(nullptr, 0, d_adds, na);
        mutate_group(m);
        mutate_apply(m);
        InvalArgs ia = inval_args(nullptr, 0, d_adds, na);
        insert_seed<W, S>(ia, false);
        relax();
        if (recover_overflow(m)) {
            insert_seed<W, S>(ia, false);
            relax();
        }
        mutate_finish(m);
    }

    void traversal_call(const cig_edge *d, uint32_t nd, const cig_edge *a, uint32_t na, bool staged,
                        cig_batch_stats *st) {
        if (uf_mode()) return uf_batch(d, nd, a, na, st);
        if (msk) {
            switch (weight_kind()) {
            case 0: msi_batch<0>(d, nd, a, na, staged, st); break;
            case 1: msi_batch<1>(d, nd, a, na, staged, st); break;
            default: msi_batch<2>(d, nd, a, na, staged, st); break;
            }
            return;
        }
        switch (weight_kind()) {
        case 0: traversal_batch<0, false>(d, nd, a, na, staged, st); break;
        case 1: traversal_batch<1, false>(d, nd, a, na, staged, st); break;
        case 2: traversal_batch<2, false>(d, nd, a, na, staged, st); break;
        default: traversal_batch<3, true>(d, nd, a, na, staged, st); break;
        }
    }

    // ------------------------------------------------ CC by union-find
    // (ccuf.hpp): labels re-derived per touched component, inserts united.
    bool uf_mode() const { return (cfg.flags & CIG_CFG_CC_UNION_REPAIR) && cfg.algo == CIG_ALGO_CC; }

    UfArgs uf_args(const cig_edge *d, uint32_t nd, const cig_edge *a, uint32_t na) {
        if (uf.count < n) {
            uf.alloc(n);
            uroot.alloc(n);
        }
        UfArgs x{};
        x.ctl = ctl.p;
        x.rows = rows.p;
        x.state = state.p;
        x.uf = uf.p;
        x.roots = uroot.p;
        x.marks = aff.p;
        x.alist = alist.p;
        x.tiles = t0.p;
        x.cfg = ecfg(kStageRepair);
        x.dels = d;
        x.nd = nd;
        x.adds = a;
        x.na = na;
        x.add_skip = cur_skip;
        x.del_skip = cur_skip ? cur_skip + na : nullptr;
        x.n = n;
        return x;
    }

    // Sample, snapshot, select, hook and flatten the listed vertices.
    template <int QPL>
    void uf_derive_qpl(const UfArgs &x) {
        k_uf_sample<<<grid_small, 256, 0, stream>>>(x);
        k_uf_snapshot<<<grid_small, 256, 0, stream>>>(x);
        k_uf_select<<<1, 256, 0, stream>>>(x);
        mark(8);
        k_uf_hook<QPL><<<grid_small, kBlock, 0, stream>>>(x);
        mark(9);
        scan_timed = stages_now;
        k_uf_hook_tiles<QPL><<<grid_small, kBlock, 0, stream>>>(x);
        k_uf_flatten<<<grid_small, 256, 0, stream>>>(x);
        HIPCHK(hipGetLastError());
    }

    void uf_derive(const UfArgs &x) {
        switch (qpl) {
        case 1: uf_derive_qpl<1>(x); break;
        case 4: uf_derive_qpl<4>(x); break;
        default: uf_derive_qpl<2>(x); break;
        }
    }

    // The initial labels: every vertex affected.
    void uf_all() {
        reset_ctl(false);  // drops the label-propagation frontier cig_init left
        UfArgs x = uf_args(nullptr, 0, nullptr, 0);
        k_uf_init<<<grid_small, 256, 0, stream>>>(x);
        uf_derive(x);
    }

    void uf_batch(const cig_edge *d, uint32_t nd, const cig_edge *a, uint32_t na, cig_batch_stats *st) {
        begin_call(st);
        MutArgs m = mut_args(d, nd, a, na);
        mutate_group(m);
        mutate_apply(m);
        mark(1);
        const UfArgs x = uf_args(d, nd, a, na);
        if (nd) {
            k_uf_mark<<<grid_small, 256, 0, stream>>>(x);
            k_uf_collect<<<grid_small, 256, 0, stream>>>(x);
            uf_derive(x);
        }
        mark(2);
        mark(3);
        if (na) {
            k_uf_union_adds<<<grid_small, 256, 0, stream>>>(x);
            k_uf_relabel<<<grid_small, 256, 0, stream>>>(x);
            HIPCHK(hipGetLastError());
        }
        mark(5);
        mark(4);
        end_call(st, 4);
        if (na && hctl.p->overflow_cnt) {  // the labels are final; only the rows need room
            recover_overflow(m);
            end_call(st, 4);
        }
        defer_finish(m, false);
    }

    // ------------------------------------------------ multi-source maintenance
    // Up to 64 BFS/SSSP sources kept current through every batch (msi.hpp):
    // the single-source steps, each run once for all sources.
    MsiArgs msi_args(const cig_edge *d, uint32_t nd, const cig_edge *a, uint32_t na) {
        MsiArgs x{};
        x.ctl = ctl.p;
        x.rows = rows.p;
        x.in_rows = in_rows_p();
        x.st = msi_st.p;
        x.act = msi_act.p;
        x.snap = msi_snap.p;
        x.affm = msi_aff.p;
        x.pinv = msi_pinv.p;
        x.stamp = stamp.p;
        x.q0 = q0.p;
        x.q1 = q1.p;
        x.t0 = t0.p;
        x.t1 = t1.p;
        x.alist = alist.p;
        x.cfg = ecfg(kStagePropagate);
        x.dels = d;
        x.nd = nd;
        x.adds = a;
        x.na = na;
        x.add_skip = cur_skip;
        x.del_skip = cur_skip ? cur_skip + na : nullptr;
        x.n = n;
        x.k = msk;
        x.seed = cfg.weight_seed;
        return x;
    }

    void msi_release() {
        msk = 0;
        ms_sources.clear();
        msi_st.release();
        msi_act.release();
        msi_snap.release();
        msi_aff.release();
        msi_pinv.release();
    }

    void msi_init(const uint32_t *sources, uint32_t k) {
        if (coop) throw HipError(CIG_EINVAL, "several sources run chained rounds: drop CIG_CFG_COOPERATIVE");
        if (!pull_rows())
            throw HipError(CIG_EINVAL, "several sources need the incoming-row index: drop CIG_CFG_SCAN_REPAIR");
        msi_release();
        msi_st.alloc(uint64_t(n) * k);
        msi_act.alloc(n);
        msi_snap.alloc(n);
        msi_aff.alloc(n);
        msi_pinv.alloc(n);
        fill32(reinterpret_cast<uint32_t *>(msi_st.p), uint64_t(n) * k * 2, kInf);
        for (auto *b : {&msi_act, &msi_snap, &msi_aff, &msi_pinv})
            fill32(reinterpret_cast<uint32_t *>(b->p), uint64_t(n) * 2, 0);
        fill32(stamp.p, n, 0);
        msk = k;
        ms_sources.assign(sources, sources + k);
        SrcList list{};
        std::memcpy(list.v, sources, size_t(k) * 4);
        k_msi_init<<<1, 64, 0, stream>>>(msi_args(nullptr, 0, nullptr, 0), list);
        HIPCHK(hipGetLastError());
    }

    template <int W, int QPL>
    void msi_relax_qpl(uint32_t grid) {
        MsiArgs a = msi_args(nullptr, 0, nullptr, 0);
        advance_round(run_chain(
            [&](uint64_t tag, uint32_t seq) {
                k_msi_consume<<<grid_small, 256, 0, stream>>>(a, a.act, prog_dev, tag, seq);
                timed_round([&] { k_msi_expand<W, false, QPL><<<grid, kBlock, 0, stream>>>(a, seq); });
                k_msi_tiles<W, false, QPL><<<grid, kBlock, 0, stream>>>(a, seq);
            },
            round_bound(), "multi-source fixed point"));
        HIPCHK(hipGetLastError());
    }

    template <int W>
    void msi_relax_w(uint32_t grid) {
        switch (qpl) {
        case 1: msi_relax_qpl<W, 1>(grid); break;
        case 4: msi_relax_qpl<W, 4>(grid); break;
        default: msi_relax_qpl<W, 2>(grid); break;
        }
    }

    void msi_relax(uint32_t grid) {
        switch (weight_kind()) {
        case 0: msi_relax_w<0>(grid); break;
        case 1: msi_relax_w<1>(grid); break;
        default: msi_relax_w<2>(grid); break;
        }
    }

    // Invalidation waves to their fixed point, then the reset that also lists
    // every affected vertex.
    template <int QPL>
    void msi_invalidate(const MsiArgs &x) {
        MsiArgs a = x;
        a.cfg = ecfg(kStageInvalidate);
        k_msi_inval_seed<<<grid_small, 256, 0, stream>>>(a);
        HIPCHK(hipGetLastError());
        advance_round(run_chain(
            [&](uint64_t tag, uint32_t seq) {
                k_msi_consume<<<grid_small, 256, 0, stream>>>(a, a.pinv, prog_dev, tag, seq);
                k_msi_expand<0, true, QPL><<<grid_round, kBlock, 0, stream>>>(a, seq);
                k_msi_tiles<0, true, QPL><<<grid_round, kBlock, 0, stream>>>(a, seq);
            },
            round_bound(), "multi-source invalidation"));
        k_msi_reset<<<grid_small, 256, 0, stream>>>(a);
        HIPCHK(hipGetLastError());
    }

    template <int W, int QPL>
    void msi_repair(const MsiArgs &x) {
        MsiArgs a = x;
        a.cfg = ecfg(kStageRepair);
        mark(8);
        k_msi_repair<W, QPL><<<grid_small, kBlock, 0, stream>>>(a);
        mark(9);
        scan_timed = stages_now;
        k_msi_repair_tiles<W, QPL><<<grid_small, kBlock, 0, stream>>>(a);
        k_msi_repair_done<<<grid_small, 256, 0, stream>>>(a);
        HIPCHK(hipGetLastError());
    }

    template <int W>
    void msi_deletes(const MsiArgs &x) {
        switch (qpl) {
        case 1: msi_invalidate<1>(x); mark(2); msi_repair<W, 1>(x); break;
        case 4: msi_invalidate<4>(x); mark(2); msi_repair<W, 4>(x); break;
        default: msi_invalidate<2>(x); mark(2); msi_repair<W, 2>(x); break;
        }
    }

    template <int W>
    void msi_insert_seed(const MsiArgs &x, bool skip_affected) {
        k_msi_insert_seed<W><<<grid_small, 256, 0, stream>>>(x, skip_affected ? 1u : 0u);
        HIPCHK(hipGetLastError());
    }

    void msi_clear() {
        k_msi_clear<<<grid_small, 256, 0, stream>>>(msi_args(nullptr, 0, nullptr, 0));
        HIPCHK(hipGetLastError());
    }

    template <int W>
    void msi_batch(const cig_edge *d, uint32_t nd, const cig_edge *a, uint32_t na, bool staged,
                   cig_batch_stats *st) {
        begin_call(st);
        if (staged && nd && na) {
            // stage 1: the deletes to their fixed point
            MutArgs m = mut_args(d, nd, nullptr, 0);
            mutate_group(m);
            mutate_apply(m);
            const MsiArgs xd = msi_args(d, nd, nullptr, 0);
            msi_deletes<W>(xd);
            msi_relax(grid_round);
            msi_clear();
            mutate_finish(m, false);
            // stage 2: the adds
            k_reset_mut<<<1, 64, 0, stream>>>(ctl.p);
            MutArgs ma = mut_args(nullptr, 0, a, na);
            mutate_group(ma);
            mutate_apply(ma);
            const MsiArgs xa = msi_args(nullptr, 0, a, na);
            msi_insert_seed<W>(xa, false);
            msi_relax(grid_round);
            mark(4);
            end_call(st, 0);
            if (hctl.p->overflow_cnt) {
                recover_overflow(ma);
                msi_insert_seed<W>(xa, false);
                msi_relax(grid_round);
                end_call(st, 0);
            }
            mutate_finish(ma);
            return;
        }
        MutArgs m = mut_args(d, nd, a, na);
        const MsiArgs x = msi_args(d, nd, a, na);
        if (nd && cfg.n_stream != 1) {
            // as traversal_batch: grouping and claims on the second stream
            // beside the invalidation waves (MsiInvalOp skips claimed slots)
            overlapped = true;
            HIPCHK(hipEventRecord(fork_ev, stream));
            HIPCHK(hipStreamWaitEvent(side, fork_ev, 0));
            mutate_group(m, side);
            const bool side_claims = claims_beside();
            if (side_claims) mutate_claims(m, side);
            HIPCHK(hipEventRecord(join_ev, side));
            try {
                switch (qpl) {
                case 1: msi_invalidate<1>(x); break;
                case 4: msi_invalidate<4>(x); break;
                default: msi_invalidate<2>(x); break;
                }
            } catch (...) {
                (void)hipStreamSynchronize(side);
                throw;
            }
            mark(2);
            HIPCHK(hipStreamWaitEvent(stream, join_ev, 0));
            mutate_apply(m, side_claims);
            mark(1);
            switch (qpl) {
            case 1: msi_repair<W, 1>(x); break;
            case 4: msi_repair<W, 4>(x); break;
            default: msi_repair<W, 2>(x); break;
            }
        } else {
            mutate_group(m);
            mutate_apply(m);
            mark(1);
            if (nd) msi_deletes<W>(x);
            else mark(2);
        }
        mark(3);
        if (na) msi_insert_seed<W>(x, nd != 0);
        mark(5);
        msi_relax(grid_round);
        mark(4);
        end_call(st, 4);
        if (na && hctl.p->overflow_cnt) {
            recover_overflow(m);
            msi_insert_seed<W>(x, false);
            msi_relax(grid_round);
            end_call(st, 4);
        }
        defer_finish(m, false);
    }

    // ------------------------------------------------------------ PR batch
    void pull_plan(const PrArgs &a, const uint32_t *gate) {
        k_pr_pull_plan_reset<<<1, 64, 0, stream>>>(ctl.p);
        k_pr_pull_plan<<<grid_small, kBlock, 0, stream>>>(a, gate);
        HIPCHK(hipGetLastError());
    }

    template <int QPL>
    void pull_pass(const PrArgs &a, uint32_t seq) {
        k_pr_pull<QPL><<<grid_pull, kBlock, 0, stream>>>(a, seq);
        k_pr_pull_long<<<grid_small, kBlock, 0, stream>>>(a, seq);
    }

    // The touched sources' contribution sign*alpha*x/deg along their current
    // rows, added to the residuals of round `round`'s pending frontier: by
    // scatter (gated by amend_push) or, with pull = true, by a sweep of the
    // incoming rows (gated by amend_pull; builds the pull plan first).
    void pr_amend(float sign, bool pull) {
        if (pull) {
            PrArgs a = pr_args(cfg.pr_max_rounds);
            fill32(reinterpret_cast<uint32_t *>(delta.p), n, 0u);
            fill32(reinterpret_cast<uint32_t *>(amask.p), 2 * ((uint64_t(n) + 63) / 64), 0u);
            k_pr_amend_delta<<<grid_small, 256, 0, stream>>>(a, touched.p, sign);
            pull_plan(a, sign < 0 ? &ctl.p->amend_pull : nullptr);
            switch (qpl) {
            case 1: pull_pass<1>(a, 0); break;
            case 4: pull_pass<4>(a, 0); break;
            default: pull_pass<2>(a, 0); break;
            }
        }
        PrScatterOp sc{rows.p, res.p, stamp.p, delta.p, &ctl.p->pr_invalid, cfg.pr_epsilon};
        PrAmendOp op{rows.p, rank.p, sign, n, sc};
        // appends to the pending frontier (queue of the next round)
        read_ctl();
        const uint32_t r = hctl.p->round;
        Sink sink{(r & 1) ? q1.p : q0.p, &ctl.p->qcount[r % 3]};
        TileSink ts{t0.p, &ctl.p->scount};
        uint32_t zero = 0;
        HIPCHK(hipMemcpyAsync(&ctl.p->scount, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        k_expand_once<1><<<grid_small, kBlock, 0, stream>>>(op, ecfg(kStageAmend), touched.p, &ctl.p->touched,
                                                            0, sink, ts, pull ? &ctl.p->amend_push : nullptr);
        k_expand_tiles<1><<<grid_small, kBlock, 0, stream>>>(op, ecfg(kStageAmend), t0.p, &ctl.p->scount, sink);
        HIPCHK(hipGetLastError());
    }

    // Engine::UpdatePageRank (framework.cuh:1660-1704):
    // r' = r + alpha (P_new^T - P_old^T) x for the touched sources, then converge.
    void pr_batch(const cig_edge *d_dels, uint32_t nd, const cig_edge *d_adds, uint32_t na,
                  cig_batch_stats *st) {
        begin_call(st);
        MutArgs m = mut_args(d_dels, nd, d_adds, na);
        mutate_group(m);
        const bool pull = pull_rows() && !coop;
        if (pull) {
            ensure_pull();
            const PrArgs a = pr_args(cfg.pr_max_rounds);
            // touched counts outgoing and incoming rows: twice the sources
            // (a symmetric graph has no separate incoming rows)
            const uint32_t min_touched = a.pull_min > 0x7FFFFFFFu ? kNone : (rev ? 2u : 1u) * a.pull_min;
            k_pr_amend_form<<<1, 64, 0, stream>>>(ctl.p, min_touched);
        }
        pr_amend(-1.0f, pull);  // old rows, old degrees
        mutate_apply(m);
        recover_overflow(m);
        pr_amend(+1.0f, pull);  // new rows, new degrees (the plan of the new rows stays for the solve)
        mark(1);
        mark(2);
        mark(3);
        mark(5);
        pr_converge(pull);
        mark(4);
        pr_residual_norm();
        mutate_finish(m);
        end_call(st, 4);
    }

    // ------------------------------------------------------------ calls
    void check_ready() {
        if (!loaded) throw HipError(CIG_ESTATE, "no graph loaded");
        if (!initialized) throw HipError(CIG_ESTATE, "cig_init must come first");
    }

    void converge(cig_batch_stats *st) {
        check_ready();
        begin_call(st);
        for (int k : {1, 2, 3, 5}) mark(k);
        if (is_pr()) pr_converge();
        else if (msk) msi_relax(grid_small);
        else if (uf_mode()) uf_all();
        else relax();
        mark(4);
        if (is_pr()) pr_residual_norm();
        end_call(st, 4);
    }

    void upload_and_apply(const cig_edge *dels, uint64_t nd, const cig_edge *adds, uint64_t na,
                          uint32_t flags, cig_batch_stats *st) {
        check_ready();
        if (nd > 0x7FFFFFFFull || na > 0x7FFFFFFFull) throw HipError(CIG_ELIMIT, "batch too large");
        if ((nd && !dels) || (na && !adds)) throw HipError(CIG_EINVAL, "null record pointer");
        ensure_records(nd, na);
        if (nd) HIPCHK(hipMemcpyAsync(rec_dels.p, dels, nd * sizeof(cig_edge), hipMemcpyHostToDevice, stream));
        if (na) HIPCHK(hipMemcpyAsync(rec_adds.p, adds, na * sizeof(cig_edge), hipMemcpyHostToDevice, stream));
        records_to_internal(rec_dels.p, nd);
        records_to_internal(rec_adds.p, na);
        apply_device(rec_dels.p, uint32_t(nd), rec_adds.p, uint32_t(na), flags, st);
    }

    // Delete tables hold at most sum(max(16, 4*nd_u)) <= 20*nd slots; scan
    // tiles are at most (live slots)/kDelTile + (rows with deletes).
    void ensure_delete_space(uint64_t nd) {
        const uint64_t tab = 20 * nd + 64;
        if (tab > 0xFFFFFFFFull) throw HipError(CIG_ELIMIT, "batch too large for the delete tables");
        if (hkeys.count < tab) {
            // table offsets are 32-bit: the growth stays below 2^32 slots
            const uint64_t n2 = std::min<uint64_t>(std::max(tab, hkeys.count + hkeys.count / 2), 0xFFFFFFFFull);
            retire(hkeys);
            retire(hcnts);
            hkeys.alloc(n2);
            hcnts.alloc(n2);
            fill32(hkeys.p, n2, 0);
            fill32(reinterpret_cast<uint32_t *>(hcnts.p), 2 * n2, 0);
        }
        grow(dtiles, (hbm_slots + host_slots) / kDelTile + nd + 64);
    }

    // Insert pieces: a relocated row gives ceil(deg/kPieceLen) copy pieces and
    // every touched source ceil(adds/kPieceLen) append pieces.
    void ensure_pieces(uint64_t nd, uint64_t na) {
        grow(pieces, (hbm_slots + host_slots) / kPieceLen + 2 * (nd + na) + na / kPieceLen + 64);
    }

    void apply_device(const cig_edge *d, uint32_t nd, const cig_edge *a, uint32_t na, uint32_t flags,
                      cig_batch_stats *st) {
        const uint64_t mirror = rev ? 2 : 1;  // the incoming rows take every update once more
        if ((uint64_t(nd) + na) * mirror > 0xFFFFFFFFull) throw HipError(CIG_ELIMIT, "batch too large");
        ensure_grouped((uint64_t(nd) + na) * mirror);
        ensure_delete_space(nd * mirror);
        ensure_pieces(nd * mirror, na * mirror);
        if (is_pr()) pr_batch(d, nd, a, na, st);
        else traversal_call(d, nd, a, na, (flags & CIG_BATCH_STAGED) != 0, st);
        if (st) st->pool_compactions = compactions;
    }

    void load_updates(const cig_edge *adds, uint64_t na, const cig_edge *dels, uint64_t nd,
                      const uint32_t *sizes, uint32_t nb) {
        if (!loaded) throw HipError(CIG_ESTATE, "no graph loaded");
        uint64_t ta = 0, td = 0, max_batch = 0;
        for (uint32_t b = 0; b < nb; ++b) {
            ta += sizes[2 * b];
            td += sizes[2 * b + 1];
            max_batch = std::max<uint64_t>(max_batch, uint64_t(sizes[2 * b]) + sizes[2 * b + 1]);
        }
        if (ta > na || td > nd) throw HipError(CIG_EINVAL, "batch sizes exceed the records");
        s_adds.alloc(ta);
        s_dels.alloc(td);
        if (ta) HIPCHK(hipMemcpy(s_adds.p, adds, ta * sizeof(cig_edge), hipMemcpyHostToDevice));
        if (td) HIPCHK(hipMemcpy(s_dels.p, dels, td * sizeof(cig_edge), hipMemcpyHostToDevice));
        records_to_internal(s_adds.p, ta);
        records_to_internal(s_dels.p, td);
        sync();
        batch_sizes.assign(sizes, sizes + 2 * nb);
        next_batch = 0;
        deletes_done = false;
        add_off = del_off = 0;
        ensure_grouped(max_batch * (rev ? 2 : 1));
    }

    void apply_next(uint32_t flags, cig_batch_stats *st) {
        check_ready();
        if (next_batch * 2 >= batch_sizes.size()) throw HipError(CIG_ELIMIT, "no update batch left");
        const uint32_t na = batch_sizes[2 * next_batch], nd = batch_sizes[2 * next_batch + 1];
        const bool del_phase = flags & CIG_BATCH_DELETE_PHASE, add_phase = flags & CIG_BATCH_ADD_PHASE;
        if (del_phase && add_phase) throw HipError(CIG_EINVAL, "choose one phase");
        if (del_phase) {
            if (deletes_done) throw HipError(CIG_ESTATE, "the deletes of this batch were applied already");
            apply_device(s_dels.p + del_off, nd, nullptr, 0, 0, st);
            deletes_done = true;
            return;
        }
        if (add_phase) {
            if (!deletes_done) throw HipError(CIG_ESTATE, "apply the delete phase of this batch first");
            apply_device(nullptr, 0, s_adds.p + add_off, na, 0, st);
        } else {
            if (deletes_done) throw HipError(CIG_ESTATE, "finish the add phase of this batch first");
            // resident records: nothing is enqueued ahead of the batch (the
            // flag is consumed by traversal_batch and dropped on any exit)
            struct Fresh {
                bool &f;
                ~Fresh() { f = false; }
            } fresh{fresh_records};
            fresh_records = true;
            apply_device(s_dels.p + del_off, nd, s_adds.p + add_off, na, flags, st);
        }
        deletes_done = false;
        add_off += na;
        del_off += nd;
        ++next_batch;
    }
    bool deletes_done = false;

    // Applies `k` consecutive update batches as one group (see k_group_match):
    // d/a are the group's deletes and adds in batch order, dcut/acut their
    // k+1 prefixes. Results equal applying the batches one by one.
    void apply_group(const cig_edge *d, uint32_t nd, const cig_edge *a, uint32_t na,
                     const std::vector<uint32_t> &dcut, const std::vector<uint32_t> &acut, uint32_t flags,
                     cig_batch_stats *st) {
        const uint32_t k = uint32_t(dcut.size() - 1);
        if (k == 1 || nd == 0 || na == 0) return apply_device(d, nd, a, na, flags, st);
        HIPCHK(hipEventRecord(ev[6], stream));
        uint32_t bits = 10;
        while ((1ull << bits) < 2ull * na) ++bits;
        const uint64_t cap = 1ull << bits, ops = uint64_t(na) + nd;
        if (gkey.count < cap) {
            gkey.alloc(cap);
            ghead.alloc(cap);
            gconf.alloc((cap + 3) & ~3ull);
        }
        if (gnext.count < ops) {
            gnext.alloc(ops);
            gskip.alloc((ops + 3) & ~3ull);
        }
        if (gcut.count < 2 * (uint64_t(k) + 1)) gcut.alloc(2 * (uint64_t(k) + 1));
        if (!gover.count) gover.alloc(1);
        fill32(reinterpret_cast<uint32_t *>(gkey.p), cap * 2, 0);
        fill32(ghead.p, cap, kNone);
        fill32(reinterpret_cast<uint32_t *>(gconf.p), (cap + 3) / 4, 0);
        fill32(reinterpret_cast<uint32_t *>(gskip.p), (ops + 3) / 4, 0);
        fill32(gover.p, 1, 0);
        std::vector<uint32_t> cuts(dcut);
        cuts.insert(cuts.end(), acut.begin(), acut.end());
        HIPCHK(hipMemcpyAsync(gcut.p, cuts.data(), cuts.size() * 4, hipMemcpyHostToDevice, stream));
        GroupArgs g{};
        g.dels = d;
        g.nd = nd;
        g.adds = a;
        g.na = na;
        g.del_cut = gcut.p;
        g.add_cut = gcut.p + k + 1;
        g.nb = k;
        g.n = n;
        g.tkey = gkey.p;
        g.thead = ghead.p;
        g.tconf = gconf.p;
        g.tbits = bits;
        g.next_op = gnext.p;
        g.skip = gskip.p;
        g.overflow = gover.p;
        g.stats = stats.p;
        k_group_adds<<<grid_small, 256, 0, stream>>>(g);
        k_group_dels<<<grid_small, 256, 0, stream>>>(g);
        k_group_match<<<grid_small, 256, 0, stream>>>(g);
        HIPCHK(hipGetLastError());
        uint32_t over = 0;
        HIPCHK(hipMemcpyAsync(&over, gover.p, 4, hipMemcpyDeviceToHost, stream));
        sync();
        if (over) {
            // a key with more than kGroupOps records: apply the batches one by
            // one (the next call's control reset drops the match statistics)
            reduce_stats();
            cig_batch_stats one{}, sum{};
            for (uint32_t b = 0; b < k; ++b) {
                apply_device(d + dcut[b], dcut[b + 1] - dcut[b], a + acut[b], acut[b + 1] - acut[b],
                             flags & ~uint32_t(CIG_BATCH_STAGED), st ? &one : nullptr);
                if (st) accumulate(sum, one);
            }
            if (st) {
                *st = sum;
                st->total_ms = elapsed(6, 7);
            }
            return;
        }
        cur_skip = gskip.p;
        try {
            apply_device(d, nd, a, na, flags & ~uint32_t(CIG_BATCH_STAGED), st);
        } catch (...) {
            cur_skip = nullptr;
            throw;
        }
        cur_skip = nullptr;
        if (st) st->total_ms = elapsed(6, 7);
    }

    static void accumulate(cig_batch_stats &sum, const cig_batch_stats &x) {
        sum.mutate_ms += x.mutate_ms;
        sum.invalidate_ms += x.invalidate_ms;
        sum.repair_ms += x.repair_ms;
        sum.seed_ms += x.seed_ms;
        sum.propagate_ms += x.propagate_ms;
        sum.edges_scanned += x.edges_scanned;
        sum.host_edges_scanned += x.host_edges_scanned;
        sum.vertices_expanded += x.vertices_expanded;
        sum.affected += x.affected;
        sum.rounds += x.rounds;
        sum.deletes_applied += x.deletes_applied;
        sum.missing_deletes += x.missing_deletes;
        sum.invalid_updates += x.invalid_updates;
        sum.adds_applied += x.adds_applied;
        sum.relocated_rows += x.relocated_rows;
        sum.touched_sources += x.touched_sources;
        sum.converged = x.converged;
        sum.pool_compactions = x.pool_compactions;
        sum.pr_residual_l1 = x.pr_residual_l1;
        for (int i = 0; i < 4; ++i) {
            sum.stage_edges[i] += x.stage_edges[i];
            sum.stage_vertices[i] += x.stage_vertices[i];
            sum.stage_host_edges[i] += x.stage_host_edges[i];
        }
        sum.scan_kernel_ms += x.scan_kernel_ms;
        sum.round_kernel_ms += x.round_kernel_ms;
        sum.round_launches += x.round_launches;
        sum.pull_rounds += x.pull_rounds;
    }

    // The next `k` resident batches as one group.
    // The next k batches one at a time (each its own call's work); st, if
    // not null, receives one entry per batch.
    void apply_next_each(uint32_t k, uint32_t flags, cig_batch_stats *st) {
        check_ready();
        if (flags & (CIG_BATCH_DELETE_PHASE | CIG_BATCH_ADD_PHASE))
            throw HipError(CIG_EINVAL, "whole batches only");
        const uint32_t left = uint32_t(batch_sizes.size() / 2) - next_batch;
        if (!k || k > left) throw HipError(CIG_ELIMIT, "not that many update batches left");
        for (uint32_t b = 0; b < k; ++b) apply_next(flags, st ? st + b : nullptr);
    }

    void apply_next_group(uint32_t k, uint32_t flags, cig_batch_stats *st) {
        check_ready();
        if (deletes_done) throw HipError(CIG_ESTATE, "finish the add phase of this batch first");
        const uint32_t left = uint32_t(batch_sizes.size() / 2) - next_batch;
        if (!k || k > left) throw HipError(CIG_ELIMIT, "not that many update batches left");
        std::vector<uint32_t> dcut(1, 0), acut(1, 0);
        for (uint32_t b = 0; b < k; ++b) {
            acut.push_back(acut.back() + batch_sizes[2 * (next_batch + b)]);
            dcut.push_back(dcut.back() + batch_sizes[2 * (next_batch + b) + 1]);
        }
        apply_group(s_dels.p + del_off, dcut.back(), s_adds.p + add_off, acut.back(), dcut, acut, flags, st);
        add_off += acut.back();
        del_off += dcut.back();
        next_batch += k;
    }

    // ------------------------------------------------------------ autotuner
    // Coordinate descent over the launch shape of the propagation kernels
    // (quads per lane, tile split threshold, tile length, relax workgroup
    // size, relax workgroups per CU), each trial timing the fixed point from
    // `src` on the loaded graph (best of two). Every shape gives the same
    // values, so only time is compared. Leaves the engine converged from src
    // with the best shape.
    void autotune(uint32_t src, cig_tune_result *res) {
        if (!loaded) throw HipError(CIG_ESTATE, "no graph loaded");
        if (is_pr() || coop) throw HipError(CIG_EINVAL, "the autotuner drives traversal engines with chained rounds");
        struct Shape {
            uint32_t qpl, split, tile, block, bpc;
        };
        auto apply = [&](const Shape &x) {
            qpl = x.qpl;
            split = x.split;
            tile_len = x.tile;
            rblock = x.block;
            grid_round = uint32_t(cus) * x.bpc;
            ensure_tiles();
        };
        auto trial = [&](const Shape &x) {
            apply(x);
            double best = 1e30;
            for (int rep = 0; rep < 2; ++rep) {
                cig_batch_stats st{};
                init(&src, 1);
                converge(&st);
                best = std::min(best, st.total_ms);
            }
            return best;
        };
        Shape cur{qpl, split, tile_len, rblock, std::max<uint32_t>(1, grid_round / uint32_t(cus))};
        const double first = trial(cur);
        double best = first;
        uint32_t trials = 1;
        const std::vector<uint32_t> qv{1, 2, 4}, sv{32, 64, 128, 256}, tv{256, 512, 1024}, bv{256, 512},
            pv{1, 2, 4};
        for (int pass = 0; pass < 2; ++pass) {
            for (int dim = 0; dim < 5; ++dim) {
                const std::vector<uint32_t> &vals = dim == 0 ? qv : dim == 1 ? sv : dim == 2 ? tv : dim == 3 ? bv : pv;
                for (uint32_t v : vals) {
                    Shape x = cur;
                    (dim == 0 ? x.qpl : dim == 1 ? x.split : dim == 2 ? x.tile : dim == 3 ? x.block : x.bpc) = v;
                    if (x.qpl == cur.qpl && x.split == cur.split && x.tile == cur.tile && x.block == cur.block &&
                        x.bpc == cur.bpc)
                        continue;
                    const double t = trial(x);
                    ++trials;
                    if (t < best) {
                        best = t;
                        cur = x;
                    }
                }
            }
        }
        apply(cur);
        init(&src, 1);
        converge(nullptr);
        if (res) {
            std::memset(res, 0, sizeof(*res));
            res->edges_per_lane = 4 * cur.qpl;
            res->split_degree = cur.split;
            res->tile_len = cur.tile;
            res->block_size = cur.block;
            res->round_blocks_per_cu = cur.bpc;
            res->trials = trials;
            res->first_ms = first;
            res->best_ms = best;
        }
    }

    // ------------------------------------------------------------ queries
    // Up to 64 BFS / SSSP sources on the live graph, independent of the
    // engine's own incremental state (which the query leaves unchanged).
    // out[s * n + v] = distance of v from sources[s] (kInf: unreachable).
    template <int W, int QPL>
    void ms_run(MsArgs &a) {
        advance_round(run_chain(
            [&](uint64_t tag, uint32_t seq) {
                // whole-graph frontiers: the one-shot grid, not the round grid
                k_ms_consume<<<grid_small, 256, 0, stream>>>(a, prog_dev, tag, seq);
                k_ms_expand<W, QPL><<<grid_small, kBlock, 0, stream>>>(a, seq);
                k_ms_tiles<W, QPL><<<grid_small, kBlock, 0, stream>>>(a, seq);
            },
            round_bound(), "multi-source query"));
    }

    template <int QPL>
    void ms_dispatch(uint32_t wkind, MsArgs &a) {
        switch (wkind) {
        case 0: ms_run<0, QPL>(a); break;
        case 1: ms_run<1, QPL>(a); break;
        default: ms_run<2, QPL>(a); break;
        }
    }

    // One caller buffer of a (possibly merged) query: the distances of its
    // `nsrc` consecutive sources.
    struct QueryOut {
        uint32_t *p;
        uint32_t nsrc;
    };

    void query(uint32_t algo, const uint32_t *sources, uint32_t k, uint32_t *out, uint64_t bytes,
               cig_batch_stats *st) {
        if (!out || bytes < uint64_t(n) * k * 4) throw HipError(CIG_EINVAL, "output buffer too small");
        const QueryOut one{out, k};
        query(algo, sources, k, &one, 1, st);
    }

    // Working memory of one multi-source run (vertex-major distances, source
    // masks, BFS reach masks, enqueue stamps).
    struct QueryMem {
        uint32_t *dist;
        unsigned long long *act, *snap, *seen;
        uint32_t *stamp;
    };

    void query_check(uint32_t algo, const uint32_t *sources, uint32_t k) const {
        if (!loaded) throw HipError(CIG_ESTATE, "no graph loaded");
        if (algo != CIG_ALGO_BFS && algo != CIG_ALGO_SSSP) throw HipError(CIG_EINVAL, "queries are BFS or SSSP");
        if (renum) throw HipError(CIG_EINVAL, "queries on a PR engine need CIG_CFG_PR_KEEP_IDS");
        if (k == 0 || k > 64 || !sources) throw HipError(CIG_EINVAL, "1..64 sources");
        for (uint32_t s = 0; s < k; ++s)
            if (sources[s] >= n) throw HipError(CIG_EINVAL, "source out of range");
    }

    // Seeds the sources and runs the rounds to the fixed point; the caller
    // has begun the call and set mem's distances to kInf and its masks to 0
    // (stamps may hold any tag below the current round, or 0).
    void query_run(uint32_t algo, const uint32_t *sources, uint32_t k, const QueryMem &mem) {
        const bool bfs = algo == CIG_ALGO_BFS;
        MsArgs a{};
        a.ctl = ctl.p;
        a.rows = rows.p;
        a.dist = mem.dist;
        a.act = mem.act;
        a.snap = mem.snap;
        a.seen = bfs ? mem.seen : nullptr;
        a.stamp = mem.stamp;
        a.q0 = q0.p;
        a.q1 = q1.p;
        a.t0 = t0.p;
        a.t1 = t1.p;
        a.cfg = ecfg(kStagePropagate);
        a.n = n;
        a.stride = (k + 3u) & ~3u;
        a.seed = cfg.weight_seed;
        SrcList list{};
        std::memcpy(list.v, sources, size_t(k) * 4);
        k_ms_seed<<<1, 64, 0, stream>>>(a, list, k);
        HIPCHK(hipGetLastError());
        for (int m : {1, 2, 3, 5}) mark(m);
        const uint32_t wkind = bfs ? 2u : (cfg.weight_mode == CIG_WEIGHT_HASH64 ? 1u : 0u);
        switch (qpl) {
        case 1: ms_dispatch<1>(wkind, a); break;
        case 4: ms_dispatch<4>(wkind, a); break;
        default: ms_dispatch<2>(wkind, a); break;
        }
        mark(4);
    }

    // The results go straight to each caller's buffer (no merged host copy).
    void query(uint32_t algo, const uint32_t *sources, uint32_t k, const QueryOut *outs, size_t nouts,
               cig_batch_stats *st) {
        query_check(algo, sources, k);
        uint64_t placed = 0;
        for (size_t i = 0; i < nouts; ++i) {
            if (!outs[i].p) throw HipError(CIG_EINVAL, "null output buffer");
            placed += outs[i].nsrc;
        }
        if (placed != k) throw HipError(CIG_EINVAL, "output buffers must cover every source");
        query_compute(algo, sources, k, st);
        query_emit(k, outs, nouts, lane);
    }

    // The rounds of a query, to its fixed point in ms_dist (vertex-major).
    // The results stay there until the next query; query_emit moves them out.
    void query_compute(uint32_t algo, const uint32_t *sources, uint32_t k, cig_batch_stats *st) {
        query_check(algo, sources, k);
        const uint32_t stride = (k + 3u) & ~3u;
        if (ms_dist.count < uint64_t(n) * stride) ms_dist.alloc(uint64_t(n) * stride);
        if (ms_act.count < n) {
            ms_act.alloc(n);
            ms_snap.alloc(n);
            ms_stamp.alloc(n);
            fill32(ms_stamp.p, n, 0);
        }
        const bool bfs = algo == CIG_ALGO_BFS;
        if (bfs && ms_seen.count < n) ms_seen.alloc(n);
        if (ms_out.count < uint64_t(n) * k) ms_out.alloc(uint64_t(n) * k);
        begin_call(st);
        fill32(ms_dist.p, uint64_t(n) * stride, kInf);
        fill32(reinterpret_cast<uint32_t *>(ms_act.p), uint64_t(n) * 2, 0);
        fill32(reinterpret_cast<uint32_t *>(ms_snap.p), uint64_t(n) * 2, 0);
        if (bfs) fill32(reinterpret_cast<uint32_t *>(ms_seen.p), uint64_t(n) * 2, 0);
        query_run(algo, sources, k, QueryMem{ms_dist.p, ms_act.p, ms_snap.p, ms_seen.p, ms_stamp.p});
        end_call(st, 4);
    }

    // Source-major results of the last query_compute of k sources into the
    // callers' buffers, on lane L's stream. Touches only ms_dist and ms_out,
    // so the serving loop runs it on its own lane while the engine stream
    // serves the next update group.
    void query_emit(uint32_t k, const QueryOut *outs, size_t nouts, CopyLane &L) {
        const uint32_t stride = (k + 3u) & ~3u;
        k_ms_gather<<<grid_small, 256, 0, L.stream>>>(ms_dist.p, n, stride, k, ms_out.p);
        HIPCHK(hipGetLastError());
        uint64_t at = 0;
        for (size_t i = 0; i < nouts; ++i) {
            L.copy(outs[i].p, ms_out.p + at, uint64_t(n) * outs[i].nsrc * 4);
            at += uint64_t(n) * outs[i].nsrc;
        }
        HIPCHK(hipStreamSynchronize(L.stream));
    }

    // Byte offsets of a run's working memory inside the no-allocation scratch
    // array: distances, act, snap, (BFS) seen, stamps, each 256-B aligned.
    static uint64_t noalloc_layout(uint32_t nv, uint32_t k, bool bfs, uint64_t off[5]) {
        auto up = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
        const uint64_t stride = (k + 3u) & ~3u;
        off[0] = 0;
        off[1] = up(off[0] + uint64_t(nv) * stride * 4);
        off[2] = up(off[1] + uint64_t(nv) * 8);
        off[3] = up(off[2] + uint64_t(nv) * 8);
        off[4] = up(off[3] + (bfs ? uint64_t(nv) * 8 : 0));
        return up(off[4] + uint64_t(nv) * 4);
    }

    // cig_query with every buffer of the run in the device's scratch array
    // (see cig_query_noalloc): no allocation on any path. The distances go
    // to the caller's device memory, source-major.
    void query_noalloc(uint32_t algo, const uint32_t *sources, uint32_t k, uint32_t *dist_dev, uint64_t bytes,
                       cig_batch_stats *st) {
        query_check(algo, sources, k);
        if (!dist_dev || bytes < uint64_t(n) * k * 4) throw HipError(CIG_EINVAL, "output buffer too small");
        const bool bfs = algo == CIG_ALGO_BFS;
        uint64_t off[5];
        const uint64_t need = noalloc_layout(n, k, bfs, off);
        if (need > CIG_NOALLOC_SCRATCH_BYTES)
            throw HipError(CIG_ELIMIT, "query needs " + std::to_string(need) + " bytes of scratch; the array holds " +
                                           std::to_string(uint64_t(CIG_NOALLOC_SCRATCH_BYTES)));
        if (!scratch_dev) throw HipError(CIG_EINTERNAL, "scratch array address not resolved");
        const uint32_t stride = (k + 3u) & ~3u;
        QueryMem m{reinterpret_cast<uint32_t *>(scratch_dev + off[0]),
                   reinterpret_cast<unsigned long long *>(scratch_dev + off[1]),
                   reinterpret_cast<unsigned long long *>(scratch_dev + off[2]),
                   bfs ? reinterpret_cast<unsigned long long *>(scratch_dev + off[3]) : nullptr,
                   reinterpret_cast<uint32_t *>(scratch_dev + off[4])};
        begin_call(st);
        // the array keeps the last call's contents (of any engine on this GPU):
        // everything the run reads before writing is cleared, stamps included
        fill32(m.dist, uint64_t(n) * stride, kInf);
        fill32(reinterpret_cast<uint32_t *>(m.act), uint64_t(n) * 2, 0);
        fill32(reinterpret_cast<uint32_t *>(m.snap), uint64_t(n) * 2, 0);
        if (bfs) fill32(reinterpret_cast<uint32_t *>(m.seen), uint64_t(n) * 2, 0);
        fill32(m.stamp, n, 0);
        query_run(algo, sources, k, m);
        k_ms_gather<<<grid_small, 256, 0, stream>>>(m.dist, n, stride, k, dist_dev);
        HIPCHK(hipGetLastError());
        end_call(st, 4);
    }

    void copy_out(void *out, const void *dev, uint64_t bytes) { lane.copy(out, dev, bytes); }

    // ------------------------------------------------------------ cache
    // Sort-based admission (compute_hot_vertices + confirm_candidate_batch,
    // framework.cuh:605-729): rows by score descending (stable: ties by id),
    // admit the longest prefix whose capacity fits 90% of the budget, evict
    // the rest to the host tier, then load admitted cold rows.
    // compute_hot_vertices + confirm_candidate_batch + evict/compact/LoadCache
    // (framework.cuh:605-769, 1873-1956) on the device: age the activity
    // window, histogram the capacities by score, pick the threshold on the
    // host from 1024 sums, list the rows that change tier, move them.
    // host_compaction: whether the evictions may compact the host tier when
    // its tail has no room (seconds at scale 27); without it such a refresh
    // moves nothing (rows stay where they are, results do not change).
    void cache_refresh(cig_batch_stats *st, bool host_compaction = true) {
        check_ready();
        begin_call(st);
        k_age_hotness<<<grid_small, 256, 0, stream>>>(hot.p, hist.p, score.p, nrows);
        HIPCHK(hipGetLastError());
        uint64_t moved = 0;
        if (budget_slots) {
            if (adm_hist.count < kScoreBuckets) {
                adm_hist.alloc(kScoreBuckets);
                adm_acc.alloc(4);
            }
            fill32(reinterpret_cast<uint32_t *>(adm_hist.p), 2 * kScoreBuckets, 0u);
            fill32(reinterpret_cast<uint32_t *>(adm_acc.p), 8, 0u);
            k_adm_hist<<<uint32_t(cus) * 2, 256, 0, stream>>>(rows.p, score.p, nrows, adm_hist.p);
            HIPCHK(hipGetLastError());
            std::vector<unsigned long long> h(kScoreBuckets);
            HIPCHK(hipMemcpyAsync(h.data(), adm_hist.p, kScoreBuckets * 8, hipMemcpyDeviceToHost, stream));
            sync();
            const uint64_t limit = budget_slots * 9 / 10;
            uint64_t above = 0;
            uint32_t thr = 0;
            for (int s = int(kScoreBuckets) - 1; s >= 0; --s) {
                if (above + h[s] > limit || s == 0) {
                    thr = uint32_t(s);
                    break;
                }
                above += h[s];
            }
            AdmArgs a{};
            a.rows = rows.p;
            a.score = score.p;
            a.nrows = nrows;
            a.thr = thr;
            a.room = limit - above;
            a.acc = adm_acc.p;
            a.evict = small_list.p;  // mutation scratch, free between calls
            a.load = big_list.p;
            a.evict_cnt = &ctl.p->small_cnt;
            a.load_cnt = &ctl.p->big_cnt;
            k_adm_mark<<<grid_small, 256, 0, stream>>>(a, 0);
            k_adm_mark<<<grid_small, 256, 0, stream>>>(a, 1);
            HIPCHK(hipGetLastError());
            unsigned long long acc[4];
            HIPCHK(hipMemcpyAsync(acc, adm_acc.p, sizeof(acc), hipMemcpyDeviceToHost, stream));
            read_ctl();
            const bool fits = host_compaction || hctl.p->pool_tail[1] + acc[1] <= hctl.p->pool_cap[1];
            if (fits) {
                if (hctl.p->small_cnt) move_rows_device(small_list.p, &ctl.p->small_cnt, hctl.p->small_cnt, 1, acc[1]);
                if (hctl.p->big_cnt) move_rows_device(big_list.p, &ctl.p->big_cnt, hctl.p->big_cnt, 0, acc[2]);
                moved = uint64_t(hctl.p->small_cnt) + hctl.p->big_cnt;
            } else {
                ++refreshes_deferred;
            }
        }
        end_call(st, 0);
        if (st) st->relocated_rows = moved;  // rows that changed tier
    }

    // Moves the `count` rows listed on the device into `tier` (`need` slots),
    // compacting that tier first when its tail has no room.
    void move_rows_device(const uint32_t *list, const uint32_t *dcount, uint32_t count, int tier, uint64_t need) {
        read_ctl();
        if (hctl.p->pool_tail[tier] + need > hctl.p->pool_cap[tier]) {
            compact_pool(tier, need);
            read_ctl();
            if (hctl.p->pool_tail[tier] + need > hctl.p->pool_cap[tier])
                throw HipError(CIG_ENOMEM, "tier pool cannot hold moved rows");
        }
        DBuf<CopyJob> jobs;
        jobs.alloc(count);
        // a row has max(1, ceil(deg / kCmpPiece)) copy pieces, deg <= cap
        DBuf<uint32_t> pjob, pcount;
        pjob.alloc(uint64_t(count) + need / kCmpPiece + 1);
        pcount.alloc(1);
        fill32(pcount.p, 1, 0u);
        k_move_plan<<<grid_small, 256, 0, stream>>>(rows.p, list, dcount, uint32_t(tier), ctl.p, jobs.p, pcount.p,
                                                    pjob.p, free_list());
        k_move_rows<<<grid_small, 256, 0, stream>>>(jobs.p, pjob.p, pcount.p, pools());
        if (tier == 0 && fl_off.p) k_move_release<<<grid_small, 256, 0, stream>>>(jobs.p, dcount, free_list());
        HIPCHK(hipGetLastError());
        sync();  // the job list is freed on return
    }

    // ------------------------------------------------------------ results
    void gather(uint32_t field, void *out, size_t bytes) {
        if (!loaded) throw HipError(CIG_ESTATE, "no graph loaded");
        sync();
        const size_t need = field == CIG_FIELD_TIER ? n : size_t(n) * 4;
        if (!out || bytes < need) throw HipError(CIG_EINVAL, "output buffer too small");
        // row / word of the caller's vertex v
        auto at = [&](uint32_t v) { return renum ? perm_h[v] : v; };
        if (field == CIG_FIELD_DEGREE || field == CIG_FIELD_TIER) {
            std::vector<Row> hrows(n);
            HIPCHK(hipMemcpy(hrows.data(), rows.p, n * sizeof(Row), hipMemcpyDeviceToHost));
            for (uint32_t v = 0; v < n; ++v) {
                const Row &r = hrows[at(v)];
                if (field == CIG_FIELD_DEGREE) static_cast<uint32_t *>(out)[v] = r.deg;
                else static_cast<uint8_t *>(out)[v] = (r.off & kHostTier) ? 0 : 1;
            }
            return;
        }
        auto words = [&](const uint32_t *src) {
            if (!renum) {
                HIPCHK(hipMemcpy(out, src, size_t(n) * 4, hipMemcpyDeviceToHost));
                return;
            }
            std::vector<uint32_t> tmp(n);
            HIPCHK(hipMemcpy(tmp.data(), src, size_t(n) * 4, hipMemcpyDeviceToHost));
            uint32_t *o = static_cast<uint32_t *>(out);
            parallel_for(n, pick_threads(0), [&](uint64_t lo, uint64_t hi, unsigned) {
                for (uint64_t v = lo; v < hi; ++v) o[v] = tmp[perm_h[v]];
            });
        };
        if (field == CIG_FIELD_HOTNESS) {
            words(score.p);
            return;
        }
        if (!initialized) throw HipError(CIG_ESTATE, "cig_init must come first");
        if (is_pr()) {
            if (field == CIG_FIELD_VALUE) words(reinterpret_cast<const uint32_t *>(rank.p));
            else if (field == CIG_FIELD_RESIDUAL) words(reinterpret_cast<const uint32_t *>(res.p));
            else throw HipError(CIG_EINVAL, "field not available for PR");
            return;
        }
        if (field != CIG_FIELD_VALUE && field != CIG_FIELD_PARENT)
            throw HipError(CIG_EINVAL, "field not available for traversal");
        if (msk) {  // source-major: out[s * n + v]
            const uint64_t words = uint64_t(n) * msk;
            if (bytes < words * 4) throw HipError(CIG_EINVAL, "output buffer too small (nsrc * n words)");
            if (ms_out.count < words) ms_out.alloc(words);
            k_msi_gather<<<grid_small, 256, 0, stream>>>(msi_st.p, n, msk, field == CIG_FIELD_VALUE ? 0u : 1u,
                                                         ms_out.p);
            HIPCHK(hipGetLastError());
            copy_out(out, ms_out.p, words * 4);
            return;
        }
        std::vector<uint64_t> hs(n);
        HIPCHK(hipMemcpy(hs.data(), state.p, size_t(n) * 8, hipMemcpyDeviceToHost));
        uint32_t *o = static_cast<uint32_t *>(out);
        for (uint32_t v = 0; v < n; ++v) o[v] = field == CIG_FIELD_VALUE ? uint32_t(hs[v] >> 32) : uint32_t(hs[v]);
    }

    // Rows [first, first+n) as CSR: the outgoing rows (first = 0) or the
    // incoming index (first = n).
    void gather_topology(uint64_t *offs, uint32_t *dst, uint64_t capacity, uint32_t first = 0) {
        if (!loaded) throw HipError(CIG_ESTATE, "no graph loaded");
        if (first && !pull_rows()) throw HipError(CIG_ESTATE, "this engine keeps no incoming-row index");
        sync();
        std::vector<Row> hrows(n);
        HIPCHK(hipMemcpy(hrows.data(), first ? in_rows_p() : rows.p, n * sizeof(Row), hipMemcpyDeviceToHost));
        read_ctl();
        const uint64_t used0 = hctl.p->pool_tail[0];
        std::vector<uint32_t> hp(std::min<uint64_t>(used0, hbm_slots));
        if (!hp.empty()) HIPCHK(hipMemcpy(hp.data(), hbm_pool.p, hp.size() * 4, hipMemcpyDeviceToHost));
        uint64_t o = 0;
        for (uint32_t v = 0; v < n; ++v) {
            offs[v] = o;
            const Row &r = hrows[renum ? perm_h[v] : v];
            if (o + r.deg > capacity) throw HipError(CIG_EINVAL, "dst capacity too small");
            const uint32_t *src = (r.off & kHostTier) ? host_pool.p + (r.off & ~kHostTier) : hp.data() + r.off;
            if (renum)
                for (uint32_t k = 0; k < r.deg; ++k) dst[o + k] = inv_h[src[k]];
            else if (r.deg)
                std::memcpy(dst + o, src, size_t(r.deg) * 4);
            o += r.deg;
        }
        offs[n] = o;
    }

    void info(cig_graph_info *gi) {
        std::memset(gi, 0, sizeof(*gi));
        gi->n = n;
        gi->algo = cfg.algo;
        gi->block_size = kBlock;
        gi->edges_per_lane = 4 * qpl;
        gi->split_degree = split;
        gi->cu_count = uint32_t(cus);
        gi->compactions = uint32_t(compactions);
        gi->inplace_compactions = uint32_t(inplace_compactions);
        gi->host_compactions = uint32_t(host_compactions);
        gi->compaction_ms_hbm = uint32_t(compaction_ms[0]);
        gi->compaction_ms_host = uint32_t(compaction_ms[1]);
        gi->refreshes_deferred = uint32_t(refreshes_deferred);
        gi->grid_blocks = grid_coop;
        if (!loaded) return;
        sync();
        std::vector<Row> hrows(n);
        HIPCHK(hipMemcpy(hrows.data(), rows.p, n * sizeof(Row), hipMemcpyDeviceToHost));
        read_ctl();
        for (const Row &r : hrows) {
            gi->edges += r.deg;
            if (!(r.off & kHostTier) && r.cap) {
                gi->cached_rows++;
                gi->cached_edges += r.deg;
            }
        }
        gi->hbm_pool_slots = hbm_slots;
        gi->hbm_pool_used = hctl.p->pool_tail[0];
        gi->host_pool_slots = host_slots;
        gi->host_pool_used = hctl.p->pool_tail[1];
        gi->host_pinned_bytes = host_slots * 4 + sizeof(Ctl);
        gi->device_bytes = hbm_slots * 4 + rows.bytes() + state.bytes() + rank.bytes() * 4 + q0.bytes() * 2 +
                           stamp.bytes() + aff.bytes() + alist.bytes() + hot.bytes() * 3 + t0.bytes() * 2 +
                           cnt64.bytes() * 3 + touched.bytes() * 5 + pieces.bytes() + grouped.bytes() +
                           rec_dels.bytes() * 2 + s_adds.bytes() + s_dels.bytes() + filt.bytes() +
                           htoff.bytes() + hcnt.bytes() + hkeys.bytes() + hcnts.bytes() + dtiles.bytes() +
                           stats.bytes();
    }

    ~cig_engine() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (side) (void)hipStreamSynchronize(side);
        for (void *p : retired) (void)hipFree(p);
        if (side) {
            (void)hipStreamSynchronize(side);
            if (fork_ev) (void)hipEventDestroy(fork_ev);
            if (join_ev) (void)hipEventDestroy(join_ev);
            if (side_done_ev) (void)hipEventDestroy(side_done_ev);
            (void)hipStreamDestroy(side);
        }
        if (stream) {
            (void)hipStreamSynchronize(stream);
            for (auto &e : ev)
                if (e) (void)hipEventDestroy(e);
            for (auto &e : rev_begin) (void)hipEventDestroy(e);
            for (auto &e : rev_end) (void)hipEventDestroy(e);
            lane.release();
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace {

// Status of the exception being handled (call inside a catch block); err
// receives its message. Used where an exception must not leave a thread.
cig_status current_exception_status(std::string &err) {
    try {
        throw;
    } catch (const HipError &x) {
        err = x.what();
        return x.code;
    } catch (const std::bad_alloc &) {
        err = "host allocation failed";
        return CIG_ENOMEM;
    } catch (const std::exception &x) {
        err = x.what();
        return CIG_EINTERNAL;
    } catch (...) {
        err = "unknown exception";
        return CIG_EINTERNAL;
    }
}

template <class F>
cig_status guarded(cig_engine *e, F f) {
    if (!e) return CIG_EINVAL;
    try {
        f();
        e->err.clear();
        return CIG_OK;
    } catch (const HipError &x) {
        e->err = x.what();
        return x.code;
    } catch (const std::bad_alloc &) {
        e->err = "host allocation failed";
        return CIG_ENOMEM;
    } catch (const std::exception &x) {
        e->err = x.what();
        return CIG_EINTERNAL;
    }
}


}  // namespace

// ------------------------------------------------------------------ serving
// One server thread owns the engine (see cig.h). Update requests are served
// in submission order, up to max_group per group; queries are served when no
// update is waiting (or once the oldest has waited max_query_wait_ms), up to
// 64 sources of one algorithm merged into one run. A query's rounds run on
// the engine stream; its results then leave the device on the server's own
// copy lane (a second stream and a host thread) while the next update group
// runs. Between groups, when nothing is queued, the cache is refreshed every
// refresh_every groups.
struct ServeRequest {
    uint64_t ticket = 0;
    uint32_t kind = 0;  // 0 update, 1 query
    std::vector<cig_edge> dels, adds;
    uint32_t algo = 0;
    std::vector<uint32_t> sources;
    uint32_t *out = nullptr;
    uint64_t out_bytes = 0;
    cig_request_info info{};
    bool done = false;
};

struct cig_server {
    cig_engine *e = nullptr;
    cig_serve_config cfg{};
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<std::shared_ptr<ServeRequest>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<ServeRequest>> live;
    uint64_t next_ticket = 1, version = 0;
    bool stopping = false;
    std::thread th;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    DBuf<cig_edge> sd, sa;  // device staging of a group's records
    CopyLane qlane;         // query results leave the device here
    std::thread emitter;    // moves the last query's results out (reads the engine's query buffers only)
    uint32_t groups_since_refresh = 0;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    uint64_t free_hbm() {
        size_t fr = 0, total = 0;
        HIPCHK(hipMemGetInfo(&fr, &total));
        return fr;
    }

    // Bytes a group of nd deletes and na adds adds to the staging buffers
    // (a buffer that grows takes at least half again its size: cig_engine::grow).
    uint64_t staging_need(uint64_t nd, uint64_t na) const {
        auto bytes = [](uint64_t want, uint64_t have) -> uint64_t {
            return want > have ? std::max(want, have + have / 2) * sizeof(cig_edge) : 0;
        };
        return bytes(nd, sd.count) + bytes(na, sa.count);
    }

    void finish(std::vector<std::shared_ptr<ServeRequest>> &batch, cig_status st, const cig_batch_stats &stats,
                uint64_t version_of_first, bool per_request_version) {
        const double t = now_ms();
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < batch.size(); ++i) {
            ServeRequest &r = *batch[i];
            r.info.status = st;
            r.info.done_ms = t;
            r.info.group_size = uint32_t(batch.size());
            r.info.version = version_of_first + (per_request_version ? i : 0);
            r.info.stats = stats;
            r.done = true;
        }
        done_cv.notify_all();
    }

    void serve_updates(std::vector<std::shared_ptr<ServeRequest>> &batch) {
        uint64_t nd = 0, na = 0;
        std::vector<uint32_t> dcut(1, 0), acut(1, 0);
        for (auto &r : batch) {
            nd += r->dels.size();
            na += r->adds.size();
            dcut.push_back(uint32_t(nd));
            acut.push_back(uint32_t(na));
        }
        cig_batch_stats stats{};
        cig_status st = CIG_OK;
        try {
            if (nd > 0x7FFFFFFFull || na > 0x7FFFFFFFull) throw HipError(CIG_ELIMIT, "group too large");
            e->grow(sd, nd);
            e->grow(sa, na);
            uint64_t od = 0, oa = 0;
            for (auto &r : batch) {
                if (!r->dels.empty())
                    HIPCHK(hipMemcpyAsync(sd.p + od, r->dels.data(), r->dels.size() * sizeof(cig_edge),
                                          hipMemcpyHostToDevice, e->stream));
                if (!r->adds.empty())
                    HIPCHK(hipMemcpyAsync(sa.p + oa, r->adds.data(), r->adds.size() * sizeof(cig_edge),
                                          hipMemcpyHostToDevice, e->stream));
                od += r->dels.size();
                oa += r->adds.size();
            }
            e->check_ready();
            e->records_to_internal(sd.p, nd);
            e->records_to_internal(sa.p, na);
            e->apply_group(sd.p, uint32_t(nd), sa.p, uint32_t(na), dcut, acut, 0, &stats);
        } catch (...) {
            st = current_exception_status(e->err);
        }
        const uint64_t v0 = version + 1;
        if (st == CIG_OK) version += batch.size();
        finish(batch, st, stats, v0, true);
        ++groups_since_refresh;
    }

    void join_emitter() {
        if (emitter.joinable()) emitter.join();
    }

    // The query's rounds on the engine stream; its results are moved out by
    // the emitter thread on the copy lane. The next query waits for the
    // emitter (it reuses the engine's query buffers), an update group does not.
    void serve_queries(std::vector<std::shared_ptr<ServeRequest>> &batch) {
        join_emitter();
        std::vector<uint32_t> src;
        for (auto &r : batch) src.insert(src.end(), r->sources.begin(), r->sources.end());
        cig_batch_stats stats{};
        try {
            e->query_compute(batch[0]->algo, src.data(), uint32_t(src.size()), &stats);
        } catch (...) {
            const cig_status st = current_exception_status(e->err);
            finish(batch, st, stats, version, false);
            return;
        }
        const uint64_t v = version;
        emitter = std::thread([this, batch, stats, v, k = uint32_t(src.size())]() mutable {
            cig_status st = CIG_OK;
            std::string err;
            try {
                (void)hipSetDevice(e->cfg.device);
                std::vector<cig_engine::QueryOut> outs;
                for (auto &r : batch) outs.push_back({r->out, uint32_t(r->sources.size())});
                e->query_emit(k, outs.data(), outs.size(), qlane);
            } catch (...) {
                st = current_exception_status(err);
            }
            finish(batch, st, stats, v, false);
        });
    }

    void fail(std::shared_ptr<ServeRequest> r, cig_status st) {
        std::vector<std::shared_ptr<ServeRequest>> one{r};
        finish(one, st, cig_batch_stats{}, version, false);
    }

    // Takes the next group from the queue (caller holds mu): the queued
    // updates in order, or the oldest query with the later ones it can merge.
    std::vector<std::shared_ptr<ServeRequest>> take_next() {
        std::vector<std::shared_ptr<ServeRequest>> batch;
        size_t first_q = queue.size(), first_u = queue.size();
        for (size_t i = 0; i < queue.size() && (first_q == queue.size() || first_u == queue.size()); ++i) {
            if (queue[i]->kind == 0 && first_u == queue.size()) first_u = i;
            if (queue[i]->kind == 1 && first_q == queue.size()) first_q = i;
        }
        const bool query_due = first_q < queue.size() &&
                               (first_u == queue.size() || now_ms() - queue[first_q]->info.submit_ms >=
                                                               double(cfg.max_query_wait_ms));
        const uint32_t kind = query_due ? 1u : 0u;
        uint32_t nsrc = 0, algo = kind ? queue[first_q]->algo : 0u;
        for (size_t i = kind ? first_q : first_u; i < queue.size();) {
            const ServeRequest &r = *queue[i];
            bool take = r.kind == kind;
            if (take && kind == 0 && batch.size() >= cfg.max_group) break;
            if (take && kind == 1) {
                take = r.algo == algo && nsrc + r.sources.size() <= cfg.max_query_sources;
                if (take) nsrc += uint32_t(r.sources.size());
            }
            if (take) {
                batch.push_back(queue[i]);
                queue.erase(queue.begin() + int64_t(i));
            } else {
                ++i;
            }
        }
        return batch;
    }

    void loop() {
        (void)hipSetDevice(e->cfg.device);
        for (;;) {
            std::vector<std::shared_ptr<ServeRequest>> batch;
            {
                std::unique_lock<std::mutex> lk(mu);
                work_cv.wait(lk, [&] { return stopping || !queue.empty(); });
                if (queue.empty()) break;  // stopping and drained
                batch = take_next();
                for (auto &r : batch) r->info.start_ms = now_ms();
            }
            if (batch.empty()) continue;
            // admission by free HBM: shrink the group until its staging fits
            if (batch[0]->kind == 0) {
                try {
                    const uint64_t fr = free_hbm();
                    uint64_t nd = 0, na = 0;
                    for (auto &r : batch) nd += r->dels.size(), na += r->adds.size();
                    while (batch.size() > 1 && staging_need(nd, na) + cfg.hbm_reserve > fr) {
                        std::lock_guard<std::mutex> lk(mu);
                        nd -= batch.back()->dels.size();
                        na -= batch.back()->adds.size();
                        queue.push_front(batch.back());
                        batch.pop_back();
                    }
                    if (staging_need(nd, na) + cfg.hbm_reserve > fr) {
                        fail(batch[0], CIG_ENOMEM);
                        continue;
                    }
                } catch (...) {
                    const cig_status st = current_exception_status(e->err);
                    for (auto &r : batch) fail(r, st);
                    continue;
                }
                serve_updates(batch);
                maybe_refresh();
            } else {
                try {
                    const uint32_t k = [&] { uint32_t t = 0; for (auto &r : batch) t += uint32_t(r->sources.size()); return t; }();
                    // distances, the source-major copy, act/snap/seen masks and stamps
                    const uint64_t need = uint64_t(e->n) * (((k + 3) & ~3u) + k) * 4 + uint64_t(e->n) * 28;
                    const uint64_t held = e->ms_dist.bytes() + e->ms_out.bytes() + e->ms_act.bytes() +
                                          e->ms_snap.bytes() + e->ms_seen.bytes() + e->ms_stamp.bytes();
                    if (need + cfg.hbm_reserve > free_hbm() + held) {
                        for (auto &r : batch) fail(r, CIG_ENOMEM);
                        continue;
                    }
                } catch (...) {
                    const cig_status st = current_exception_status(e->err);
                    for (auto &r : batch) fail(r, st);
                    continue;
                }
                serve_queries(batch);
            }
        }
        join_emitter();
    }

    // Cache maintenance between update groups (the reference refreshes after
    // every batch, hybrid_sssp.cu:474-498): every refresh_every groups, when
    // no request is waiting. Results do not depend on it.
    void maybe_refresh() {
        if (!cfg.refresh_every || groups_since_refresh < cfg.refresh_every || !e->budget_slots) return;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (!queue.empty()) return;
        }
        groups_since_refresh = 0;
        try {
            e->cache_refresh(nullptr, false);  // never a host-tier compaction between groups
            ++refreshes;
        } catch (...) {
            std::string err;
            (void)current_exception_status(err);
            refresh_error = err;
        }
    }
    uint64_t refreshes = 0;
    std::string refresh_error;
};

extern "C" {

cig_status cig_config_default(cig_config *cfg) {
    if (!cfg) return CIG_EINVAL;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->abi_version = CIG_ABI_VERSION;
    cfg->algo = CIG_ALGO_SSSP;
    cfg->weight_mode = CIG_WEIGHT_REF128;
    cfg->n_stream = 3;
    cfg->segments = 1;
    cfg->pr_max_rounds = 100;
    cfg->pr_epsilon = 1e-6f;
    return CIG_OK;
}

cig_status cig_create(const cig_config *cfg, cig_engine **out) {
    g_create_error.clear();
    if (!cfg || !out) { g_create_error = "null argument"; return CIG_EINVAL; }
    *out = nullptr;
    if (cfg->abi_version != CIG_ABI_VERSION) { g_create_error = "ABI version mismatch"; return CIG_EINVAL; }
    if (cfg->algo > CIG_ALGO_PR || cfg->weight_mode > CIG_WEIGHT_HASH64) {
        g_create_error = "unknown algo or weight mode";
        return CIG_EINVAL;
    }
    if (cfg->algo == CIG_ALGO_PR &&
        (!(cfg->pr_epsilon > 0.0f) || !std::isfinite(cfg->pr_epsilon) || cfg->pr_max_rounds < 1 ||
         cfg->pr_max_rounds > 1000000)) {
        g_create_error = "PR requires finite epsilon > 0 and 1 <= pr_max_rounds <= 1000000";
        return CIG_EINVAL;
    }
    if (cfg->segments > 512) { g_create_error = "segments must be <= 512"; return CIG_EINVAL; }
    if ((cfg->flags & CIG_CFG_CC_UNION_REPAIR) &&
        (cfg->algo != CIG_ALGO_CC || !(cfg->flags & CIG_CFG_SYMMETRIC) || (cfg->flags & CIG_CFG_COOPERATIVE))) {
        g_create_error = "CIG_CFG_CC_UNION_REPAIR needs CC with CIG_CFG_SYMMETRIC and chained rounds";
        return CIG_EINVAL;
    }
    std::unique_ptr<cig_engine> e(new (std::nothrow) cig_engine());
    if (!e) return CIG_ENOMEM;
    e->cfg = *cfg;
    try {
        e->setup_device();
    } catch (const HipError &x) {
        g_create_error = x.what();
        return x.code;
    }
    *out = e.release();
    return CIG_OK;
}

void cig_destroy(cig_engine *e) {
    if (!e) return;
    if (e->server) (void)cig_serve_stop(e->server);  // its thread uses the engine until it stops
    delete e;
}

const char *cig_last_error(const cig_engine *e) { return e ? e->err.c_str() : g_create_error.c_str(); }

cig_status cig_load_csr(cig_engine *e, uint32_t n, const uint64_t *offsets, const uint32_t *dst) {
    return guarded(e, [&] {
        if (!offsets || (!dst && n && offsets[n])) throw HipError(CIG_EINVAL, "null CSR pointer");
        e->load(n, offsets, dst);
    });
}

cig_status cig_graph_info_get(cig_engine *e, cig_graph_info *info) {
    return guarded(e, [&] {
        if (!info) throw HipError(CIG_EINVAL, "null info");
        e->info(info);
    });
}

cig_status cig_gather_topology(cig_engine *e, uint64_t *offsets, uint32_t *dst, uint64_t cap) {
    return guarded(e, [&] {
        if (!offsets || (!dst && cap)) throw HipError(CIG_EINVAL, "null output");
        e->gather_topology(offsets, dst, cap);
    });
}

cig_status cig_gather_incoming(cig_engine *e, uint64_t *offsets, uint32_t *src, uint64_t cap) {
    return guarded(e, [&] {
        if (!offsets || (!src && cap)) throw HipError(CIG_EINVAL, "null output");
        e->gather_topology(offsets, src, cap, e->n);
    });
}

cig_status cig_init(cig_engine *e, const uint32_t *sources, uint32_t nsrc) {
    return guarded(e, [&] { e->init(sources, nsrc); });
}

cig_status cig_converge(cig_engine *e, cig_batch_stats *st) {
    return guarded(e, [&] { e->converge(st); });
}

cig_status cig_delete_edges(cig_engine *e, const cig_edge *dels, uint64_t nd, cig_batch_stats *st) {
    return guarded(e, [&] { e->upload_and_apply(dels, nd, nullptr, 0, 0, st); });
}

cig_status cig_insert_edges(cig_engine *e, const cig_edge *adds, uint64_t na, cig_batch_stats *st) {
    return guarded(e, [&] { e->upload_and_apply(nullptr, 0, adds, na, 0, st); });
}

cig_status cig_apply_batch(cig_engine *e, const cig_edge *dels, uint64_t nd, const cig_edge *adds,
                           uint64_t na, uint32_t flags, cig_batch_stats *st) {
    return guarded(e, [&] { e->upload_and_apply(dels, nd, adds, na, flags, st); });
}

cig_status cig_load_updates(cig_engine *e, const cig_edge *adds, uint64_t na, const cig_edge *dels,
                            uint64_t nd, const uint32_t *sizes, uint32_t nb) {
    return guarded(e, [&] {
        if ((!sizes && nb) || (!adds && na) || (!dels && nd)) throw HipError(CIG_EINVAL, "null argument");
        e->load_updates(adds, na, dels, nd, sizes, nb);
    });
}

cig_status cig_autotune(cig_engine *e, uint32_t source, cig_tune_result *res) {
    return guarded(e, [&] {
        if (source >= e->n && e->loaded) throw HipError(CIG_EINVAL, "source out of range");
        e->autotune(source, res);
    });
}

cig_status cig_query(cig_engine *e, uint32_t algo, const uint32_t *sources, uint32_t nsrc, uint32_t *dist,
                     uint64_t dist_bytes, cig_batch_stats *st) {
    return guarded(e, [&] { e->query(algo, sources, nsrc, dist, dist_bytes, st); });
}

cig_status cig_query_noalloc(cig_engine *e, uint32_t algo, const uint32_t *sources, uint32_t nsrc,
                             uint32_t *dist_device, uint64_t dist_bytes, cig_batch_stats *st) {
    return guarded(e, [&] { e->query_noalloc(algo, sources, nsrc, dist_device, dist_bytes, st); });
}

cig_status cig_apply_next_batches(cig_engine *e, uint32_t count, uint32_t flags, cig_batch_stats *st) {
    return guarded(e, [&] {
        if (flags & CIG_BATCHES_EACH) e->apply_next_each(count, flags & ~uint32_t(CIG_BATCHES_EACH), st);
        else e->apply_next_group(count, flags, st);
    });
}

cig_status cig_apply_next_batch(cig_engine *e, uint32_t flags, cig_batch_stats *st) {
    return guarded(e, [&] { e->apply_next(flags, st); });
}

uint32_t cig_update_batches_left(const cig_engine *e) {
    if (!e) return 0;
    const uint32_t total = uint32_t(e->batch_sizes.size() / 2);
    return total > e->next_batch ? total - e->next_batch : 0;
}

cig_status cig_cache_refresh(cig_engine *e, cig_batch_stats *st) {
    return guarded(e, [&] { e->cache_refresh(st); });
}

cig_status cig_compact(cig_engine *e, uint32_t tier) {
    return guarded(e, [&] {
        if (tier > 1) throw HipError(CIG_EINVAL, "tier is 0 (HBM) or 1 (host)");
        if (!e->loaded) throw HipError(CIG_ESTATE, "no graph loaded");
        e->flush_finish();
        e->read_ctl();
        if (e->hctl.p->pool_tail[tier] == 0) return;  // nothing placed in this tier
        e->compact_pool(int(tier), 0);
        e->sync();
    });
}

cig_status cig_gather(cig_engine *e, uint32_t field, void *out, size_t bytes) {
    return guarded(e, [&] { e->gather(field, out, bytes); });
}

cig_status cig_copy_probe(int32_t device, uint64_t bytes, uint32_t iters, double *ms) {
    if (!ms || bytes < 16 || !iters) return CIG_EINVAL;
    try {
        HIPCHK(hipSetDevice(device));
        const uint64_t n16 = bytes / 16;
        DBuf<uint4> a, b;
        a.alloc(n16);
        b.alloc(n16);
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        const uint32_t grid = uint32_t(prop.multiProcessorCount) * 16;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        k_copy_probe<<<grid, 256>>>(a.p, b.p, n16);  // warm
        HIPCHK(hipEventRecord(e0, nullptr));
        for (uint32_t i = 0; i < iters; ++i) k_copy_probe<<<grid, 256>>>(a.p, b.p, n16);
        HIPCHK(hipEventRecord(e1, nullptr));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = double(t) / iters;
        return CIG_OK;
    } catch (const HipError &x) {
        g_create_error = x.what();
        return x.code;
    }
}

cig_status cig_host_read_probe(int32_t device, uint64_t bytes, uint32_t iters, double *ms) {
    if (!ms || bytes < 16 || !iters) return CIG_EINVAL;
    try {
        HIPCHK(hipSetDevice(device));
        const uint64_t n16 = bytes / 16;
        HBuf<uint4> h;
        h.alloc(n16);
        std::memset(h.p, 0x5A, n16 * 16);
        DBuf<uint32_t> sink;
        sink.alloc(1);
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        const uint32_t grid = uint32_t(prop.multiProcessorCount) * 16;
        const uint4 *src = h.device_ptr();
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        k_host_read_probe<<<grid, 256>>>(src, n16, sink.p);  // warm
        HIPCHK(hipEventRecord(e0, nullptr));
        for (uint32_t i = 0; i < iters; ++i) k_host_read_probe<<<grid, 256>>>(src, n16, sink.p);
        HIPCHK(hipEventRecord(e1, nullptr));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = double(t) / iters;
        return CIG_OK;
    } catch (const HipError &x) {
        g_create_error = x.what();
        return x.code;
    }
}


cig_status cig_serve_config_default(cig_serve_config *c) {
    if (!c) return CIG_EINVAL;
    std::memset(c, 0, sizeof(*c));
    c->max_group = 32;
    c->max_query_sources = 64;
    c->hbm_reserve = uint64_t(1) << 30;
    c->refresh_every = 8;
    c->max_query_wait_ms = 250;
    return CIG_OK;
}

cig_status cig_serve_start(cig_engine *e, const cig_serve_config *c, cig_server **out) {
    if (!e || !out) return CIG_EINVAL;
    return guarded(e, [&] {
        e->check_ready();
        cig_serve_config cfg;
        cig_serve_config_default(&cfg);
        if (c) cfg = *c;
        if (!cfg.max_group || !cfg.max_query_sources || cfg.max_query_sources > 64)
            throw HipError(CIG_EINVAL, "max_group >= 1 and max_query_sources in 1..64");
        if (e->server) throw HipError(CIG_ESTATE, "a server already runs on this engine");
        std::unique_ptr<cig_server> sp(new cig_server);
        sp->e = e;
        sp->cfg = cfg;
        sp->qlane.init(nullptr);  // a stream of its own
        cig_server *s = sp.release();
        s->th = std::thread([s] { s->loop(); });
        e->server = s;
        *out = s;
    });
}

static cig_status submit(cig_server *s, std::shared_ptr<ServeRequest> r, uint64_t *ticket) {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return CIG_ESTATE;
    r->ticket = s->next_ticket++;
    r->info.status = CIG_EPENDING;
    r->info.submit_ms = s->now_ms();
    s->live[r->ticket] = r;
    s->queue.push_back(r);
    *ticket = r->ticket;
    s->work_cv.notify_one();
    return CIG_OK;
}

cig_status cig_submit_update(cig_server *s, const cig_edge *dels, uint64_t nd, const cig_edge *adds, uint64_t na,
                             uint64_t *ticket) {
    if (!s || !ticket || (nd && !dels) || (na && !adds)) return CIG_EINVAL;
    try {
        auto r = std::make_shared<ServeRequest>();
        r->kind = 0;
        r->info.kind = 0;
        r->dels.assign(dels, dels + nd);
        r->adds.assign(adds, adds + na);
        return submit(s, r, ticket);
    } catch (const std::bad_alloc &) {
        return CIG_ENOMEM;
    }
}

cig_status cig_submit_query(cig_server *s, uint32_t algo, const uint32_t *sources, uint32_t nsrc, uint32_t *dist,
                            uint64_t dist_bytes, uint64_t *ticket) {
    if (!s || !ticket || !sources || !dist || nsrc == 0 || nsrc > s->cfg.max_query_sources) return CIG_EINVAL;
    if (algo != CIG_ALGO_BFS && algo != CIG_ALGO_SSSP) return CIG_EINVAL;
    if (dist_bytes < uint64_t(s->e->n) * nsrc * 4) return CIG_EINVAL;
    for (uint32_t i = 0; i < nsrc; ++i)
        if (sources[i] >= s->e->n) return CIG_EINVAL;
    try {
        auto r = std::make_shared<ServeRequest>();
        r->kind = 1;
        r->info.kind = 1;
        r->algo = algo;
        r->sources.assign(sources, sources + nsrc);
        r->out = dist;
        r->out_bytes = dist_bytes;
        return submit(s, r, ticket);
    } catch (const std::bad_alloc &) {
        return CIG_ENOMEM;
    }
}

static cig_status collect(cig_server *s, uint64_t ticket, cig_request_info *info, bool block) {
    if (!s || !info) return CIG_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    std::shared_ptr<ServeRequest> r;
    {
        auto it = s->live.find(ticket);
        if (it == s->live.end()) return CIG_EINVAL;
        r = it->second;  // no iterator is kept across the wait: submissions may rehash `live`
    }
    if (block) s->done_cv.wait(lk, [&] { return r->done; });
    if (!r->done) return CIG_EPENDING;
    // Another waiter on the same ticket may have collected it meanwhile.
    if (s->live.erase(ticket) == 0) return CIG_EINVAL;
    *info = r->info;
    return CIG_OK;
}

cig_status cig_poll(cig_server *s, uint64_t ticket, cig_request_info *info) { return collect(s, ticket, info, false); }
cig_status cig_wait(cig_server *s, uint64_t ticket, cig_request_info *info) { return collect(s, ticket, info, true); }

cig_status cig_serve_stop(cig_server *s) {
    if (!s) return CIG_EINVAL;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->work_cv.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    if (s->e && s->e->server == s) s->e->server = nullptr;
    delete s;
    return CIG_OK;
}

int cig_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

}  // extern "C"
