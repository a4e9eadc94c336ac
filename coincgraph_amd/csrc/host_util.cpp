// Host-side utilities of libcig: R-MAT graph and update-stream synthesis and
// the reference's text formats. No GPU code here.
//
// R-MAT follows the quadrant recursion of the GAP generator the reference
// vendors (include/framework/generator.h:80-113: a,b,c = 0.57,0.19,0.19,
// d = 1-a-b-c), with a counter-based RNG so that any thread count yields the
// same graph, and without the id permutation (vertex 0 stays the hub).
// Formats: market_big graph text (src/utils/parser.cpp:292-402) and the
// update/batch-size files (include/framework/Loader.h:28-99).
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cig.h"
#include "common.h"
#include "host_par.h"

namespace cig {

uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

namespace {

// One R-MAT sample: `scale` quadrant choices from 24-bit uniforms drawn from
// a counter-based stream keyed by (seed, sample index).
struct RmatSampler {
    uint32_t scale;
    double ab, a, abc;
    uint64_t seed;
    inline void sample(uint64_t idx, uint32_t &src, uint32_t &dst) const {
        uint64_t state = splitmix64(seed ^ splitmix64(idx * 0xd1b54a32d192ed03ull + 1));
        uint32_t s = 0, d = 0;
        uint64_t bits = 0;
        for (uint32_t depth = 0; depth < scale; ++depth) {
            if ((depth & 1) == 0) { state += 0x9e3779b97f4a7c15ull; bits = splitmix64(state); }
            const uint32_t r24 = (depth & 1) ? uint32_t(bits >> 40) : uint32_t((bits >> 8) & 0xFFFFFF);
            const double r = r24 * (1.0 / 16777216.0);
            s <<= 1;
            d <<= 1;
            if (r < ab) {
                if (r > a) d |= 1;
            } else {
                s |= 1;
                if (r > abc) d |= 1;
            }
        }
        src = s;
        dst = d;
    }
};

// Sorts 64-bit keys that use only the low `key_bits` bits: one parallel MSD
// partition on the top 11 of those bits, then every bucket is sorted on its own
// (buckets fit in cache; threads take buckets from a shared counter).
void radix_sort_u64(std::vector<uint64_t> &keys, unsigned key_bits, unsigned threads) {
    const uint64_t n = keys.size();
    if (n < 2) return;
    const unsigned kDigit = 11, kBuckets = 1u << kDigit;
    const unsigned shift = key_bits > kDigit ? key_bits - kDigit : 0;
    threads = std::max(1u, std::min<unsigned>(threads, unsigned((n + 65535) / 65536)));
    std::vector<uint64_t> tmp(n);
    std::vector<uint64_t> hist(size_t(threads) * kBuckets, 0);
    const uint64_t chunk = (n + threads - 1) / threads;
    auto bucket = [&](uint64_t k) { return unsigned((k >> shift) & (kBuckets - 1)); };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            uint64_t *h = &hist[size_t(t) * kBuckets];
            for (uint64_t i = std::min(n, t * chunk); i < std::min(n, (t + 1) * chunk); ++i) h[bucket(keys[i])]++;
        });
    for (auto &th : pool) th.join();
    pool.clear();
    std::vector<uint64_t> start(kBuckets + 1, 0);
    uint64_t total = 0;
    for (unsigned bkt = 0; bkt < kBuckets; ++bkt) {
        start[bkt] = total;
        for (unsigned t = 0; t < threads; ++t) {
            const uint64_t c = hist[size_t(t) * kBuckets + bkt];
            hist[size_t(t) * kBuckets + bkt] = total;
            total += c;
        }
    }
    start[kBuckets] = total;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            uint64_t *h = &hist[size_t(t) * kBuckets];
            for (uint64_t i = std::min(n, t * chunk); i < std::min(n, (t + 1) * chunk); ++i)
                tmp[h[bucket(keys[i])]++] = keys[i];
        });
    for (auto &th : pool) th.join();
    pool.clear();
    std::atomic<unsigned> next{0};
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (unsigned bkt; (bkt = next.fetch_add(1)) < kBuckets;)
                std::sort(tmp.begin() + int64_t(start[bkt]), tmp.begin() + int64_t(start[bkt + 1]));
        });
    for (auto &th : pool) th.join();
    keys.swap(tmp);
}

// Open-addressing set of 64-bit edge keys with a position payload, used by
// the stream generator for uniform delete sampling with swap-remove.
class EdgeIndex {
public:
    explicit EdgeIndex(uint64_t expected) {
        uint64_t cap = 1024;
        while (cap < expected * 2) cap <<= 1;
        keys_.assign(cap, kEmpty);
        vals_.assign(cap, 0);
        mask_ = cap - 1;
    }
    bool contains(uint64_t k) const { return find_slot(k) != kNone; }
    void insert(uint64_t k, uint64_t pos) {
        uint64_t i = hash(k) & mask_;
        while (keys_[i] != kEmpty) i = (i + 1) & mask_;
        keys_[i] = k;
        vals_[i] = pos;
    }
    uint64_t get(uint64_t k) const { return vals_[find_slot(k)]; }
    void set(uint64_t k, uint64_t pos) { vals_[find_slot(k)] = pos; }
    void erase(uint64_t k) {  // backward-shift deletion
        uint64_t i = find_slot(k);
        if (i == kNone) return;
        uint64_t j = i;
        for (;;) {
            j = (j + 1) & mask_;
            if (keys_[j] == kEmpty) break;
            uint64_t home = hash(keys_[j]) & mask_;
            bool move = (j > i) ? (home <= i || home > j) : (home <= i && home > j);
            if (move) {
                keys_[i] = keys_[j];
                vals_[i] = vals_[j];
                i = j;
            }
        }
        keys_[i] = kEmpty;
    }

private:
    static constexpr uint64_t kEmpty = ~0ull, kNone = ~0ull;
    static uint64_t hash(uint64_t k) { return splitmix64(k); }
    uint64_t find_slot(uint64_t k) const {
        uint64_t i = hash(k) & mask_;
        while (keys_[i] != kEmpty) {
            if (keys_[i] == k) return i;
            i = (i + 1) & mask_;
        }
        return kNone;
    }
    std::vector<uint64_t> keys_, vals_;
    uint64_t mask_;
};

template <class T>
T *malloc_copy(const std::vector<T> &v) {
    T *p = static_cast<T *>(std::malloc(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace
}  // namespace cig

using namespace cig;

extern "C" cig_status cig_rmat_csr(uint32_t scale, uint32_t edge_factor, double a, double b,
                                   double c, uint64_t seed, uint32_t symmetric, uint32_t threads,
                                   uint32_t *n_out, uint64_t **offsets_out, uint32_t **dst_out) {
    if (!n_out || !offsets_out || !dst_out || scale == 0 || scale > 31 || edge_factor == 0 ||
        a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return CIG_EINVAL;
    const unsigned T = pick_threads(threads);
    const uint32_t n = 1u << scale;
    const uint64_t samples = uint64_t(edge_factor) << scale;
    RmatSampler smp{scale, a + b, a, a + b + c, seed};
    // packed key (src << scale | dst) sorts on 2*scale bits; every self-loop is
    // mapped to the largest key, (n-1, n-1), itself a self-loop, dropped below
    const uint64_t self_key = (uint64_t(n - 1) << scale) | (n - 1);
    std::vector<uint64_t> keys;
    try {
        keys.resize(symmetric ? samples * 2 : samples);
    } catch (...) {
        return CIG_ENOMEM;
    }
    parallel_for(samples, T, [&](uint64_t lo, uint64_t hi, unsigned) {
        for (uint64_t i = lo; i < hi; ++i) {
            uint32_t s, d;
            smp.sample(i, s, d);
            const bool loop = s == d;
            if (symmetric) {
                keys[2 * i] = loop ? self_key : (uint64_t(s) << scale | d);
                keys[2 * i + 1] = loop ? self_key : (uint64_t(d) << scale | s);
            } else {
                keys[i] = loop ? self_key : (uint64_t(s) << scale | d);
            }
        }
    });
    radix_sort_u64(keys, 2 * scale, T);
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if (!keys.empty() && keys.back() == self_key) keys.pop_back();
    uint64_t *off = static_cast<uint64_t *>(std::calloc(uint64_t(n) + 1, sizeof(uint64_t)));
    uint32_t *dst = static_cast<uint32_t *>(std::malloc(std::max<uint64_t>(1, keys.size()) * sizeof(uint32_t)));
    if (!off || !dst) { std::free(off); std::free(dst); return CIG_ENOMEM; }
    const uint64_t mask = n - 1;
    for (uint64_t k : keys) off[(k >> scale) + 1]++;
    for (uint32_t i = 0; i < n; ++i) off[i + 1] += off[i];
    parallel_for(keys.size(), T, [&](uint64_t lo, uint64_t hi, unsigned) {
        for (uint64_t i = lo; i < hi; ++i) dst[i] = uint32_t(keys[i] & mask);
    });
    *n_out = n;
    *offsets_out = off;
    *dst_out = dst;
    return CIG_OK;
}

namespace {

// Live edge set of the stream generator: the initial CSR (rows sorted, each
// occurrence deletable through a bitset) plus the edges inserted so far.
class LiveEdges {
public:
    LiveEdges(uint32_t n, const uint64_t *off, const uint32_t *dst, uint64_t grow)
        : n_(n), off_(off, off + n + 1), dst_(dst, dst + off[n]), src_(off[n]),
          dead_((off[n] + 63) / 64, 0), ins_index_(grow + 16) {
        const unsigned T = pick_threads(0);
        parallel_for(n, T, [&](uint64_t lo, uint64_t hi, unsigned) {
            for (uint64_t s = lo; s < hi; ++s) {  // rows must be sorted for lookups
                uint32_t *b = dst_.data() + off_[s], *e = dst_.data() + off_[s + 1];
                if (!std::is_sorted(b, e)) std::sort(b, e);
                std::fill(src_.begin() + int64_t(off_[s]), src_.begin() + int64_t(off_[s + 1]), uint32_t(s));
            }
        });
        live_ = off[n];
        ins_.reserve(grow);
        ins_dead_.reserve(grow);
    }
    uint64_t live() const { return live_; }
    // a live initial occurrence of (u,v), or m() if none
    uint64_t find_initial(uint32_t u, uint32_t v) const {
        const uint32_t *b = dst_.data() + off_[u], *e = dst_.data() + off_[u + 1];
        for (const uint32_t *p = std::lower_bound(b, e, v); p != e && *p == v; ++p) {
            const uint64_t i = uint64_t(p - dst_.data());
            if (!is_dead(i)) return i;
        }
        return m();
    }
    bool contains(uint32_t u, uint32_t v) const {
        return find_initial(u, v) != m() || ins_index_.contains(key(u, v));
    }
    void insert(uint32_t u, uint32_t v) {
        ins_index_.insert(key(u, v), ins_.size());
        ins_.push_back(key(u, v));
        ins_dead_.push_back(0);
        ++live_;
    }
    // removes one live occurrence of (u,v); false if there is none
    bool remove(uint32_t u, uint32_t v) {
        const uint64_t i = find_initial(u, v);
        if (i != m()) { kill(i); return true; }
        const uint64_t k = key(u, v);
        if (!ins_index_.contains(k)) return false;
        ins_dead_[ins_index_.get(k)] = 1;
        ins_index_.erase(k);
        --live_;
        return true;
    }
    // uniform over live edges by rejection over all slots ever created
    bool sample(uint64_t &rng, uint32_t &u, uint32_t &v) {
        if (!live_) return false;
        const uint64_t slots = m() + ins_.size();
        for (;;) {
            rng += 0x9e3779b97f4a7c15ull;
            const uint64_t i = splitmix64(rng) % slots;
            if (i < m()) {
                if (is_dead(i)) continue;
                u = src_[i];
                v = dst_[i];
                kill(i);
                return true;
            }
            const uint64_t j = i - m();
            if (ins_dead_[j]) continue;
            u = uint32_t(ins_[j] >> 32);
            v = uint32_t(ins_[j]);
            ins_dead_[j] = 1;
            ins_index_.erase(ins_[j]);
            --live_;
            return true;
        }
    }

private:
    static uint64_t key(uint32_t u, uint32_t v) { return uint64_t(u) << 32 | v; }
    uint64_t m() const { return dst_.size(); }
    bool is_dead(uint64_t i) const { return (dead_[i >> 6] >> (i & 63)) & 1; }
    void kill(uint64_t i) { dead_[i >> 6] |= 1ull << (i & 63); --live_; }
    uint32_t n_;
    std::vector<uint64_t> off_;
    std::vector<uint32_t> dst_;
    std::vector<uint32_t> src_;
    std::vector<uint64_t> dead_;
    std::vector<uint64_t> ins_;
    std::vector<uint8_t> ins_dead_;
    EdgeIndex ins_index_;
    uint64_t live_ = 0;
};

}  // namespace

extern "C" cig_status cig_rmat_stream(uint32_t n, const uint64_t *offsets, const uint32_t *dst,
                                      uint32_t scale, double a, double b, double c,
                                      uint64_t seed, uint32_t nbatches, uint64_t n_del,
                                      uint64_t n_add, uint32_t symmetric, cig_edge **adds_out,
                                      uint64_t *n_add_out, cig_edge **dels_out,
                                      uint64_t *n_del_out, uint32_t **batch_sizes_out) {
    if (!offsets || (!dst && offsets[n] != 0) || !adds_out || !dels_out || !batch_sizes_out ||
        !n_add_out || !n_del_out || scale > 31 || n == 0 || (1ull << scale) > uint64_t(n) * 2)
        return CIG_EINVAL;
    const uint64_t grow = uint64_t(nbatches) * n_add * (symmetric ? 2 : 1);
    std::vector<cig_edge> adds, dels;
    std::vector<uint32_t> sizes;
    try {
        LiveEdges live(n, offsets, dst, grow);
        adds.reserve(grow);
        dels.reserve(uint64_t(nbatches) * n_del * (symmetric ? 2 : 1));
        RmatSampler smp{scale, a + b, a, a + b + c, splitmix64(seed ^ 0x5eed5eed5eedull)};
        uint64_t rng = splitmix64(seed ^ 0xde1e7edull), draw = 0;
        for (uint32_t bt = 0; bt < nbatches; ++bt) {
            uint64_t nd = 0;
            for (uint64_t i = 0; i < n_del; ++i) {
                uint32_t u, v;
                if (!live.sample(rng, u, v)) break;
                dels.push_back({u, v, 1});
                ++nd;
                if (symmetric) {
                    live.remove(v, u);
                    dels.push_back({v, u, 1});
                    ++nd;
                }
            }
            uint64_t na = 0, attempts = 0;
            while (na < n_add * (symmetric ? 2 : 1)) {
                if (++attempts > 64 * (n_add + 16)) return CIG_ELIMIT;  // graph saturated
                uint32_t u, v;
                smp.sample(draw++, u, v);
                if (u == v || u >= n || v >= n) continue;
                if (live.contains(u, v) || (symmetric && live.contains(v, u))) continue;
                live.insert(u, v);
                adds.push_back({u, v, 1});
                ++na;
                if (symmetric) {
                    live.insert(v, u);
                    adds.push_back({v, u, 1});
                    ++na;
                }
            }
            sizes.push_back(uint32_t(na));
            sizes.push_back(uint32_t(nd));
        }
    } catch (...) {
        return CIG_ENOMEM;
    }
    *adds_out = malloc_copy(adds);
    *dels_out = malloc_copy(dels);
    *batch_sizes_out = malloc_copy(sizes);
    if (!*adds_out || !*dels_out || !*batch_sizes_out) return CIG_ENOMEM;
    *n_add_out = adds.size();
    *n_del_out = dels.size();
    return CIG_OK;
}

namespace {

bool read_file(const char *path, std::string &out) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(size > 0 ? size_t(size) : 0);
    bool ok = size <= 0 || std::fread(&out[0], 1, size_t(size), f) == size_t(size);
    std::fclose(f);
    return ok;
}

// Parses up to `max_fields` unsigned integers from one line starting at p.
const char *parse_line(const char *p, const char *end, uint64_t *vals, int max_fields, int &count,
                       char *tag) {
    count = 0;
    if (tag) *tag = 0;
    while (p < end && *p != '\n') {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
        if (p >= end || *p == '\n') break;
        if (*p >= '0' && *p <= '9') {
            uint64_t v = 0;
            while (p < end && *p >= '0' && *p <= '9') v = v * 10 + uint64_t(*p++ - '0');
            if (count < max_fields) vals[count] = v;
            ++count;
        } else {
            // a non-numeric token is only allowed as the leading record tag
            if (tag && count == 0 && !*tag) *tag = *p;
            else count = -1000000;
            while (p < end && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') ++p;
        }
    }
    return p < end ? p + 1 : p;
}

}  // namespace

extern "C" cig_status cig_read_market(const char *path, uint32_t *n_out, uint64_t **offsets_out,
                                      uint32_t **dst_out) {
    if (!path || !n_out || !offsets_out || !dst_out) return CIG_EINVAL;
    std::string text;
    if (!read_file(path, text)) return CIG_EIO;
    std::vector<uint32_t> src, dstv;
    uint64_t maxid = 0;
    bool any = false;
    const char *p = text.data(), *end = p + text.size();
    while (p < end) {
        uint64_t vals[3];
        int cnt;
        p = parse_line(p, end, vals, 3, cnt, nullptr);
        if (cnt == 0) continue;
        if (cnt < 2 || cnt > 3 || vals[0] >= 0xFFFFFFFFull || vals[1] >= 0xFFFFFFFFull) return CIG_EIO;
        src.push_back(uint32_t(vals[0]));
        dstv.push_back(uint32_t(vals[1]));
        maxid = std::max<uint64_t>(maxid, std::max(vals[0], vals[1]));
        any = true;
    }
    const uint32_t n = any ? uint32_t(maxid + 1) : 0;
    uint64_t *off = static_cast<uint64_t *>(std::calloc(uint64_t(n) + 1, sizeof(uint64_t)));
    uint32_t *dst = static_cast<uint32_t *>(std::malloc(std::max<size_t>(1, dstv.size()) * sizeof(uint32_t)));
    if (!off || !dst) { std::free(off); std::free(dst); return CIG_ENOMEM; }
    for (uint32_t s : src) off[s + 1]++;
    for (uint32_t i = 0; i < n; ++i) off[i + 1] += off[i];
    std::vector<uint64_t> fill(off, off + n + (n ? 0 : 1));
    for (size_t i = 0; i < src.size(); ++i) dst[fill[src[i]]++] = dstv[i];  // file order per row
    *n_out = n;
    *offsets_out = off;
    *dst_out = dst;
    return CIG_OK;
}

extern "C" cig_status cig_read_updates(const char *update_path, const char *batch_path,
                                       cig_edge **adds_out, uint64_t *n_add_out,
                                       cig_edge **dels_out, uint64_t *n_del_out,
                                       uint32_t **batch_sizes_out, uint32_t *nbatches_out) {
    if (!update_path || !batch_path || !adds_out || !dels_out || !batch_sizes_out ||
        !n_add_out || !n_del_out || !nbatches_out)
        return CIG_EINVAL;
    std::string text;
    if (!read_file(batch_path, text)) return CIG_EIO;
    std::vector<uint32_t> sizes;
    uint64_t total_add = 0, total_del = 0;
    const char *p = text.data(), *end = p + text.size();
    while (p < end) {
        uint64_t vals[2];
        int cnt;
        p = parse_line(p, end, vals, 2, cnt, nullptr);
        if (cnt == 0) continue;
        if (cnt != 2) return CIG_EIO;
        sizes.push_back(uint32_t(vals[0]));
        sizes.push_back(uint32_t(vals[1]));
        total_add += vals[0];
        total_del += vals[1];
    }
    if (!read_file(update_path, text)) return CIG_EIO;
    std::vector<cig_edge> adds, dels;
    p = text.data();
    end = p + text.size();
    while (p < end) {
        uint64_t vals[3] = {0, 0, 1};
        int cnt;
        char tag;
        p = parse_line(p, end, vals, 3, cnt, &tag);
        if (cnt == 0 && !tag) continue;
        if (cnt < 2 || cnt > 3 || (tag != 'a' && tag != 'd')) return CIG_EIO;
        cig_edge e{uint32_t(vals[0]), uint32_t(vals[1]), uint32_t(cnt >= 3 ? vals[2] : 1)};
        (tag == 'a' ? adds : dels).push_back(e);
    }
    if (adds.size() < total_add || dels.size() < total_del) return CIG_EIO;
    *adds_out = malloc_copy(adds);
    *dels_out = malloc_copy(dels);
    *batch_sizes_out = malloc_copy(sizes);
    if (!*adds_out || !*dels_out || !*batch_sizes_out) return CIG_ENOMEM;
    *n_add_out = adds.size();
    *n_del_out = dels.size();
    *nbatches_out = uint32_t(sizes.size() / 2);
    return CIG_OK;
}

// ------------------------------------------------------------ binary formats
// Little-endian, a 32-byte header then the arrays exactly as the C-ABI
// holds them, so that reading is one sequential read per array:
//   graph  "CIGG" v1: n (u32), m (u64), offsets[n+1] (u64), dst[m] (u32)
//   stream "CIGU" v1: nbatches (u32), n_add (u64), n_del (u64),
//                     sizes[2*nbatches] (u32: adds, dels per batch),
//                     adds[n_add], dels[n_del] (cig_edge, 12 B)
namespace {

constexpr uint32_t kGraphMagic = 0x47474943u;   // "CIGG"
constexpr uint32_t kStreamMagic = 0x55474943u;  // "CIGU"
constexpr uint32_t kBinVersion = 1;

struct BinHeader {
    uint32_t magic, version, count32, pad;
    uint64_t a, b;
};
static_assert(sizeof(BinHeader) == 32, "binary header layout");

struct File {
    FILE *f = nullptr;
    explicit File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
    bool write(const void *p, size_t bytes) { return !bytes || std::fwrite(p, 1, bytes, f) == bytes; }
    bool read(void *p, size_t bytes) { return !bytes || std::fread(p, 1, bytes, f) == bytes; }
    bool at_end() { return std::fgetc(f) == EOF; }
    // bytes between the current position and the end of the file
    uint64_t remaining() {
        const long pos = std::ftell(f);
        if (pos < 0 || std::fseek(f, 0, SEEK_END) != 0) return 0;
        const long end = std::ftell(f);
        std::fseek(f, pos, SEEK_SET);
        return end > pos ? uint64_t(end - pos) : 0;
    }
};

template <class T>
T *read_array(File &in, uint64_t count, bool &ok) {
    // a header count is untrusted: count * sizeof(T) must not wrap, and the
    // file must hold that many bytes after the current position
    if (count > SIZE_MAX / sizeof(T) || count * sizeof(T) > in.remaining()) {
        ok = false;
        return nullptr;
    }
    T *p = static_cast<T *>(std::malloc(std::max<uint64_t>(1, count) * sizeof(T)));
    if (!p) return nullptr;
    ok = ok && in.read(p, count * sizeof(T));
    return p;
}

}  // namespace

extern "C" cig_status cig_write_csr_bin(const char *path, uint32_t n, const uint64_t *offsets, const uint32_t *dst) {
    if (!path || !offsets || (offsets[n] && !dst)) return CIG_EINVAL;
    File out(path, "wb");
    if (!out.f) return CIG_EIO;
    const BinHeader h{kGraphMagic, kBinVersion, n, 0, offsets[n], 0};
    const bool ok = out.write(&h, sizeof(h)) && out.write(offsets, (uint64_t(n) + 1) * 8) &&
                    out.write(dst, offsets[n] * 4);
    return ok ? CIG_OK : CIG_EIO;
}

// Out-of-core R-MAT: the samples of cig_rmat_csr (same counter-based
// sampler and seed, so the same graph), written to a CIGG file one range of
// source segments at a time. Each pass regenerates every sample and keeps the
// edges whose source lies in its range (and, undirected, those whose
// destination does, reversed); a pass covers consecutive segments whose
// expected edge count fits `mem_bytes`. All edges of a source fall in one
// segment, so dedup and the per-row sort are local to a pass; nothing of size
// m is held in memory.
extern "C" cig_status cig_rmat_write(const char *path, uint32_t scale, uint32_t edge_factor, double a, double b,
                                     double c, uint64_t seed, uint32_t symmetric, uint32_t segments,
                                     uint64_t mem_bytes, uint32_t threads, uint64_t *m_out) {
    if (!path || scale == 0 || scale > 31 || edge_factor == 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return CIG_EINVAL;
    const uint32_t n = 1u << scale;
    if (segments == 0) segments = std::min<uint32_t>(512, n);
    if (segments > n || (segments & (segments - 1))) return CIG_EINVAL;
    unsigned sbits = 0;
    while ((1u << sbits) < segments) ++sbits;
    const unsigned T = pick_threads(threads);
    const uint64_t samples = uint64_t(edge_factor) << scale;
    const RmatSampler smp{scale, a + b, a, a + b + c, seed};
    const uint32_t seg_len = n >> sbits;
    // expected edges of segment p: its source prefix (a+b for a 0 bit, c+d for
    // a 1 bit, top bit first), plus for an undirected graph its destination
    // prefix (a+c, b+d)
    std::vector<double> expect(segments);
    for (uint32_t p = 0; p < segments; ++p) {
        double ps = 1, pd = 1;
        for (unsigned k = 0; k < sbits; ++k) {
            const bool one = (p >> (sbits - 1 - k)) & 1;
            ps *= one ? 1.0 - (a + b) : a + b;
            pd *= one ? 1.0 - (a + c) : a + c;
        }
        expect[p] = double(samples) * (ps + (symmetric ? pd : 0.0));
    }
    const double per_pass = std::max(1.0, double(mem_bytes) / 24.0);  // keys, sort buffer, thread parts
    File out(path, "wb");
    if (!out.f) return CIG_EIO;
    const uint64_t dst_base = sizeof(BinHeader) + (uint64_t(n) + 1) * 8;
    uint64_t m = 0;
    try {
        for (uint32_t p0 = 0; p0 < segments;) {
            uint32_t p1 = p0 + 1;
            double e = expect[p0];
            while (p1 < segments && e + expect[p1] <= per_pass) e += expect[p1++];
            const uint32_t lo = p0 * seg_len, hi = p1 == segments ? n : p1 * seg_len;
            std::vector<std::vector<uint64_t>> part(T);
            parallel_for(samples, T, [&](uint64_t i0, uint64_t i1, unsigned t) {
                std::vector<uint64_t> &v = part[t];
                for (uint64_t i = i0; i < i1; ++i) {
                    uint32_t su, dv;
                    smp.sample(i, su, dv);
                    if (su == dv) continue;
                    if (su >= lo && su < hi) v.push_back(uint64_t(su) << scale | dv);
                    if (symmetric && dv >= lo && dv < hi) v.push_back(uint64_t(dv) << scale | su);
                }
            });
            uint64_t total = 0;
            for (auto &v : part) total += v.size();
            std::vector<uint64_t> keys;
            keys.reserve(total);
            for (auto &v : part) {
                keys.insert(keys.end(), v.begin(), v.end());
                std::vector<uint64_t>().swap(v);
            }
            radix_sort_u64(keys, 2 * scale, T);
            keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
            std::vector<uint64_t> off(size_t(hi - lo) + 1, 0);
            for (uint64_t k : keys) off[(k >> scale) - lo + 1]++;
            off[0] = m;
            for (uint32_t i = 0; i < hi - lo; ++i) off[i + 1] += off[i];
            std::vector<uint32_t> dst(keys.size());
            for (size_t i = 0; i < keys.size(); ++i) dst[i] = uint32_t(keys[i] & (n - 1));
            if (std::fseek(out.f, long(sizeof(BinHeader) + uint64_t(lo) * 8), SEEK_SET) != 0 ||
                !out.write(off.data(), (uint64_t(hi) - lo) * 8))
                return CIG_EIO;
            if (std::fseek(out.f, long(dst_base + m * 4), SEEK_SET) != 0 || !out.write(dst.data(), dst.size() * 4))
                return CIG_EIO;
            m += keys.size();
            p0 = p1;
        }
    } catch (const std::bad_alloc &) {
        return CIG_ENOMEM;
    }
    const BinHeader h{kGraphMagic, kBinVersion, n, 0, m, 0};
    if (std::fseek(out.f, long(sizeof(BinHeader) + uint64_t(n) * 8), SEEK_SET) != 0 || !out.write(&m, 8) ||
        std::fseek(out.f, 0, SEEK_SET) != 0 || !out.write(&h, sizeof(h)))
        return CIG_EIO;
    if (m_out) *m_out = m;
    return CIG_OK;
}

extern "C" cig_status cig_read_csr_bin(const char *path, uint32_t *n_out, uint64_t **offsets_out,
                                       uint32_t **dst_out) {
    if (!path || !n_out || !offsets_out || !dst_out) return CIG_EINVAL;
    File in(path, "rb");
    if (!in.f) return CIG_EIO;
    BinHeader h{};
    if (!in.read(&h, sizeof(h)) || h.magic != kGraphMagic || h.version != kBinVersion) return CIG_EIO;
    bool ok = true;
    uint64_t *off = read_array<uint64_t>(in, uint64_t(h.count32) + 1, ok);
    if (!off) return ok ? CIG_ENOMEM : CIG_EIO;
    ok = ok && off[0] == 0 && off[h.count32] == h.a;
    for (uint32_t v = 0; ok && v < h.count32; ++v) ok = off[v] <= off[v + 1];
    uint32_t *dst = ok ? read_array<uint32_t>(in, h.a, ok) : nullptr;
    if (ok && !dst) {
        std::free(off);
        return CIG_ENOMEM;
    }
    if (!ok || !in.at_end()) {
        std::free(off);
        std::free(dst);
        return CIG_EIO;
    }
    *n_out = h.count32;
    *offsets_out = off;
    *dst_out = dst;
    return CIG_OK;
}

extern "C" cig_status cig_write_updates_bin(const char *path, const cig_edge *adds, uint64_t n_add,
                                            const cig_edge *dels, uint64_t n_del, const uint32_t *batch_sizes,
                                            uint32_t nbatches) {
    if (!path || (n_add && !adds) || (n_del && !dels) || (nbatches && !batch_sizes)) return CIG_EINVAL;
    File out(path, "wb");
    if (!out.f) return CIG_EIO;
    const BinHeader h{kStreamMagic, kBinVersion, nbatches, 0, n_add, n_del};
    const bool ok = out.write(&h, sizeof(h)) && out.write(batch_sizes, uint64_t(nbatches) * 8) &&
                    out.write(adds, n_add * sizeof(cig_edge)) && out.write(dels, n_del * sizeof(cig_edge));
    return ok ? CIG_OK : CIG_EIO;
}

extern "C" cig_status cig_read_updates_bin(const char *path, cig_edge **adds_out, uint64_t *n_add_out,
                                           cig_edge **dels_out, uint64_t *n_del_out, uint32_t **batch_sizes_out,
                                           uint32_t *nbatches_out) {
    if (!path || !adds_out || !n_add_out || !dels_out || !n_del_out || !batch_sizes_out || !nbatches_out)
        return CIG_EINVAL;
    File in(path, "rb");
    if (!in.f) return CIG_EIO;
    BinHeader h{};
    if (!in.read(&h, sizeof(h)) || h.magic != kStreamMagic || h.version != kBinVersion) return CIG_EIO;
    bool ok = true;
    uint32_t *sizes = read_array<uint32_t>(in, uint64_t(h.count32) * 2, ok);
    cig_edge *adds = read_array<cig_edge>(in, h.a, ok);
    cig_edge *dels = read_array<cig_edge>(in, h.b, ok);
    if (!sizes || !adds || !dels) {
        std::free(sizes);
        std::free(adds);
        std::free(dels);
        return ok ? CIG_ENOMEM : CIG_EIO;
    }
    uint64_t ta = 0, td = 0;
    for (uint32_t b = 0; ok && b < h.count32; ++b) {
        ta += sizes[2 * b];
        td += sizes[2 * b + 1];
    }
    // the batch sizes may not ask for more records than the file holds
    // (the text reader's rule, Loader.h:28-99)
    if (!ok || ta > h.a || td > h.b || !in.at_end()) {
        std::free(sizes);
        std::free(adds);
        std::free(dels);
        return CIG_EIO;
    }
    *adds_out = adds;
    *n_add_out = h.a;
    *dels_out = dels;
    *n_del_out = h.b;
    *batch_sizes_out = sizes;
    *nbatches_out = h.count32;
    return CIG_OK;
}

extern "C" void cig_host_free(void *p) { std::free(p); }

extern "C" uint32_t cig_edge_weight(uint32_t weight_mode, uint64_t seed, uint32_t u, uint32_t v) {
    return weight_mode == CIG_WEIGHT_HASH64
               ? uint32_t(1u + splitmix64(seed ^ (uint64_t(u) << 32 | v)) % 64u)
               : uint32_t((uint64_t(u) + v) % 128u + 1u);
}
