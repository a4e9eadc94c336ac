// Shared host/device declarations of libcig.
#pragma once
#include <cstddef>
#include <cstdint>

namespace cig {

uint64_t splitmix64(uint64_t x);

// Per-vertex row descriptor of the two-tier segmented CSR (16 B, one
// dwordx4 load). `off` is a slot offset into the HBM pool, or into the pinned
// host pool when bit 63 is set. Rows own `cap` slots; `deg` are live.
struct Row {
    uint64_t off;
    uint32_t deg;
    uint32_t cap;
};
static_assert(sizeof(Row) == 16, "Row must stay 16 bytes");

constexpr uint64_t kHostTier = 1ull << 63;
constexpr uint32_t kInf = 0xFFFFFFFFu;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMinRowCap = 4;  // RequiredChunkCapacity minimum (source_local_chunk_store.h:131-144)

inline uint64_t next_pow2_cap(uint64_t deg) {
    if (deg == 0) return 0;
    uint64_t c = deg < kMinRowCap ? kMinRowCap : deg;
    --c;
    c |= c >> 1; c |= c >> 2; c |= c >> 4; c |= c >> 8; c |= c >> 16; c |= c >> 32;
    return c + 1;
}

// One tile of a long row, expanded by one wave.
struct Tile {
    uint64_t start;   // pool offset (tier bit included) of the first slot
    uint32_t v;       // row owner
    uint32_t len;     // slots in this tile
    uint32_t payload; // op payload (PR contribution bits); unused by traversal
    uint32_t pad;
};
static_assert(sizeof(Tile) == 24, "Tile layout");

// Copy job for relocating a long row to a new block.
struct CopyJob {
    uint64_t from;  // pool offset with tier bit
    uint64_t to;
    uint32_t len;
    uint32_t pfirst;  // first of the job's copy pieces (k_move_rows)
    uint32_t cap;     // block capacity (slots)
    uint32_t pad;
};

// Device control block: counters of the current batch and of the fixed-point
// rounds. Lives in device memory; zeroed per batch by a fill kernel. The
// counters that many workgroups of one kernel advance with atomics each
// start a 128-B line of their own (kCtlLine): atomics to one line are
// serialised where the line lives, and a workgroup step of k_plan advances
// eight or nine cursors at once.
constexpr size_t kCtlLine = 128;
struct Ctl {
    // fixed-point rounds (relax / PR); slot = absolute round % 3
    alignas(kCtlLine) uint32_t qcount[3];
    alignas(kCtlLine) uint32_t tcount[3];
    uint32_t round;        // next absolute round to run
    uint32_t converged;
    alignas(kCtlLine) uint32_t done;  // workgroups finished in the current round kernel
    uint32_t iround;       // invalidation: first round of the running chain
    uint32_t istart[3];    // invalidation: where round slot k's frontier starts in the affected list
    uint32_t ifront[3];    // invalidation: frontier length of round slot k
    // invalidation
    alignas(kCtlLine) uint32_t acount;  // affected list length
    alignas(kCtlLine) uint32_t icount[3];  // invalidation tile counts
    uint32_t scount;       // one-shot scan tile count
    // mutation
    alignas(kCtlLine) uint32_t touched;
    alignas(kCtlLine) uint32_t small_cnt;
    alignas(kCtlLine) uint32_t big_cnt;
    alignas(kCtlLine) uint32_t add_cnt;
    alignas(kCtlLine) uint32_t copy_cnt;
    alignas(kCtlLine) uint32_t overflow_cnt;
    alignas(kCtlLine) uint32_t mid_cnt;       // medium rows with deletes: HBM tier
    alignas(kCtlLine) uint32_t mid_host_cnt;  // and host tier
    alignas(kCtlLine) uint32_t mid16_cnt;     // HBM medium rows with 9-16 deletes
    uint32_t mid_pad;
    alignas(kCtlLine) uint32_t dtile_cnt;     // long-row delete tiles
    alignas(kCtlLine) uint32_t ht_cursor;     // and their table slots
    uint32_t pr_invalid;
    alignas(kCtlLine) uint32_t group_cursor;  // records grouped so far (< 2^32 per call)
    uint32_t pl_tiles;      // PR pull plan: tiles of long incoming rows
    uint32_t pl_multi;      // PR pull plan: rows of more than one tile
    uint32_t pull_rounds;   // PR rounds of the current call run in pull form
    uint32_t amend_pull;    // PR batch amendment gathers (1) or scatters (amend_push = 1)
    uint32_t amend_push;
    uint32_t bmin[3];       // delta-ordered rounds: smallest bucket of round slot k's frontier
    uint32_t ufsel[2];      // union-find repair: the sampled component left out of hooking, its votes
    alignas(kCtlLine) uint64_t pool_tail[2];  // bump tails of the HBM (0) and host (1) pools
    uint64_t pool_cap[2];
    uint64_t garbage[2];   // slots abandoned by relocations
    // statistics of the current call
    uint64_t edges_scanned;
    uint64_t host_edges;
    uint64_t vexp;
    uint64_t deletes_applied;
    uint64_t missing;
    uint64_t invalid;
    uint64_t adds_applied;
    uint64_t relocated;
    uint64_t touched_sources;
    uint64_t rounds;
    double residual_l1;
    // per expansion pass: 0 invalidation, 1 repair scan, 2 propagation, 3 PR amend
    uint64_t st_edges[4];
    uint64_t st_vexp[4];
    uint64_t st_host[4];
};

// Statistics are added into one of kStatShards 256-B shards (picked by the
// adding wave) so that no single address takes every wave's atomic; a
// one-workgroup kernel folds the shards into the control block.
constexpr uint32_t kStatShards = 64;
struct StatShard {
    uint64_t edges[4], vexp[4], host[4];  // per expansion pass (Stage)
    uint64_t deletes_applied, missing, invalid, adds_applied, relocated;
    uint64_t garbage[2];
    uint64_t touched;      // touched sources (outgoing rows only)
    double l1;             // PR residual norm
    uint64_t pad[11];
};
static_assert(sizeof(StatShard) == 256, "StatShard layout");
constexpr uint32_t kStatWords = 20;  // leading uint64 fields of a shard

enum Stage : uint32_t { kStageInvalidate = 0, kStageRepair = 1, kStagePropagate = 2, kStageAmend = 3 };

}  // namespace cig
