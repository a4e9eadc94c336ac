// --check verification of the samples: the reference's runtime self-checks,
// restated over the topology the engine reports back (cig_gather_topology).
//   CheckBellmanRelaxed / CheckParentWitness / DistanceChecksum:
//     samples/hybrid_sssp/hybrid_sssp.cu:211-330 (BFS: unit weights)
//   CCCheck / LabelChecksum: samples/hybrid_cc/hybrid_cc_host.cu:8-32,
//     hybrid_cc.cu:189-201
//   PageRankCheck: samples/hybrid_pr/hybrid_pr_host.cu:9-56
#pragma once
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <functional>
#include <vector>

namespace cigsample {

struct Csr {
    uint32_t n = 0;
    std::vector<uint64_t> off;
    std::vector<uint32_t> dst;
};

struct BellmanResult {
    uint64_t relaxable_edges = 0, missing_tight_witnesses = 0;
    uint32_t first_src = UINT32_MAX, first_dst = UINT32_MAX, first_missing_witness = UINT32_MAX;
};

struct ParentResult {
    uint64_t invalid_vertices = 0, missing_parent_edges = 0, wrong_parent_distances = 0;
    uint32_t first_vertex = UINT32_MAX, first_parent = UINT32_MAX;
};

using WeightFn = std::function<uint64_t(uint32_t, uint32_t)>;

inline BellmanResult check_bellman(const Csr &g, const std::vector<uint32_t> &dist, uint32_t src,
                                   const WeightFn &w) {
    BellmanResult r;
    std::vector<uint8_t> tight(g.n, 0);
    if (src < g.n) tight[src] = 1;
    for (uint32_t s = 0; s < g.n; ++s) {
        if (dist[s] == UINT32_MAX) continue;
        for (uint64_t e = g.off[s]; e < g.off[s + 1]; ++e) {
            const uint32_t d = g.dst[e];
            const uint64_t c = uint64_t(dist[s]) + w(s, d);
            if (dist[d] != UINT32_MAX && c == dist[d]) tight[d] = 1;
            if (c < dist[d]) {
                if (!r.relaxable_edges) { r.first_src = s; r.first_dst = d; }
                ++r.relaxable_edges;
            }
        }
    }
    for (uint32_t v = 0; v < g.n; ++v) {
        if (dist[v] == UINT32_MAX || tight[v]) continue;
        if (!r.missing_tight_witnesses) r.first_missing_witness = v;
        ++r.missing_tight_witnesses;
    }
    return r;
}

inline ParentResult check_parent(const Csr &g, const std::vector<uint32_t> &dist,
                                 const std::vector<uint32_t> &parent, uint32_t src,
                                 const WeightFn &w) {
    ParentResult r;
    for (uint32_t v = 0; v < g.n; ++v) {
        if (v == src || dist[v] == UINT32_MAX) continue;
        const uint32_t p = parent[v];
        bool valid = p < g.n && dist[p] != UINT32_MAX;
        if (valid) {
            bool exists = false;
            for (uint64_t e = g.off[p]; e < g.off[p + 1] && !exists; ++e) exists = g.dst[e] == v;
            const bool matches = uint64_t(dist[p]) + w(p, v) == dist[v];
            if (!exists) ++r.missing_parent_edges;
            else if (!matches) ++r.wrong_parent_distances;
            valid = exists && matches;
        }
        if (!valid) {
            if (!r.invalid_vertices) { r.first_vertex = v; r.first_parent = p; }
            ++r.invalid_vertices;
        }
    }
    return r;
}

inline uint64_t distance_checksum(const std::vector<uint32_t> &d) {
    uint64_t c = 1469598103934665603ull;
    for (uint32_t v = 0; v < d.size(); ++v) {
        if (d[v] == UINT32_MAX) continue;
        c ^= uint64_t(v) + 0x9e3779b97f4a7c15ull + (uint64_t(d[v]) << 6) + (uint64_t(d[v]) >> 2);
        c *= 1099511628211ull;
    }
    return c;
}

inline uint64_t parent_checksum(const std::vector<uint32_t> &d, const std::vector<uint32_t> &p) {
    uint64_t c = 1099511628211ull;
    for (uint32_t v = 0; v < d.size(); ++v) {
        if (d[v] == UINT32_MAX) continue;
        c ^= uint64_t(v) + 0x517cc1b727220a95ull + (uint64_t(p[v]) << 7) + (uint64_t(p[v]) >> 3);
        c *= 1469598103934665603ull;
    }
    return c;
}

// Flood from seeds in ascending id order: expected[v] = min seed reaching v.
inline uint64_t cc_check(const Csr &g, const std::vector<uint32_t> &labels) {
    std::vector<uint32_t> expected(g.n, UINT32_MAX), stack;
    for (uint32_t seed = 0; seed < g.n; ++seed) {
        if (expected[seed] != UINT32_MAX) continue;
        expected[seed] = seed;
        stack.push_back(seed);
        while (!stack.empty()) {
            const uint32_t u = stack.back();
            stack.pop_back();
            for (uint64_t e = g.off[u]; e < g.off[u + 1]; ++e) {
                const uint32_t v = g.dst[e];
                if (expected[v] != UINT32_MAX) continue;
                expected[v] = seed;
                stack.push_back(v);
            }
        }
    }
    uint64_t errors = 0;
    for (uint32_t u = 0; u < g.n; ++u) errors += labels[u] != expected[u];
    return errors;
}

struct PrCheck {
    bool ok = false, converged = false, finite = true;
    double l1 = 0, bound = 0, max_error = 0, defect = 0, max_residual = 0;
};

inline void pr_multiply(const Csr &g, const std::vector<double> &x, std::vector<double> &next) {
    std::fill(next.begin(), next.end(), 0.15);
    for (uint32_t s = 0; s < g.n; ++s) {
        const uint64_t deg = g.off[s + 1] - g.off[s];
        if (!deg) continue;
        const double c = 0.85 * x[s] / double(deg);
        for (uint64_t e = g.off[s]; e < g.off[s + 1]; ++e) next[g.dst[e]] += c;
    }
}

inline PrCheck pr_check(const Csr &g, const std::vector<float> &ranks,
                        const std::vector<float> &residual, double epsilon) {
    PrCheck r;
    const size_t n = g.n;
    std::vector<double> ref(n, 0.0), next(n), actual(ranks.begin(), ranks.end());
    for (unsigned round = 0; round < 10000; ++round) {
        pr_multiply(g, ref, next);
        double delta = 0;
        for (size_t v = 0; v < n; ++v) delta += std::abs(next[v] - ref[v]);
        ref.swap(next);
        if (delta <= 1e-12 * double(std::max<size_t>(1, n))) { r.converged = true; break; }
    }
    pr_multiply(g, actual, next);
    double mass = 0;
    for (size_t v = 0; v < n; ++v) {
        r.finite = r.finite && std::isfinite(ranks[v]) && std::isfinite(residual[v]) && ranks[v] >= -epsilon;
        const double err = std::abs(double(ranks[v]) - ref[v]);
        r.l1 += err;
        r.max_error = std::max(r.max_error, err);
        r.defect += std::abs(next[v] - ranks[v] - residual[v]);
        r.max_residual = std::max(r.max_residual, double(std::abs(residual[v])));
        mass += std::abs(ref[v]);
    }
    const double roundoff = 64 * double(FLT_EPSILON) * std::max(1.0, mass);
    r.bound = (double(n) * epsilon + roundoff) / 0.15;
    r.ok = r.finite && r.converged && r.l1 <= r.bound && r.defect <= roundoff &&
           r.max_residual <= double(float(epsilon));
    return r;
}

}  // namespace cigsample
