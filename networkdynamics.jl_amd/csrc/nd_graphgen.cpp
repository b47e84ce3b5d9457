// Seeded synthetic graphs for the workloads (Erdos-Renyi G(n,m), Barabasi-Albert,
// random geometric, Watts-Strogatz).  These are this library's own generators:
// Graphs.jl's random streams cannot be reproduced without Julia.  Output edge
// lists are simple (no self-loops, no duplicates) and sorted like Graphs.jl's
// `edges(g)` of a SimpleGraph: (src < dst), lexicographic.
//
// PRNG: xoshiro256** (Blackman & Vigna) seeded through splitmix64; bounded
// integers by Lemire's multiply-shift with rejection; doubles from the top 53 bits.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <new>
#include <unordered_set>
#include <vector>

#include "nd_internal.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) {
            x += 0x9E3779B97F4A7C15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    uint64_t below(uint64_t n) {   // uniform in [0, n)
        uint64_t x = next();
        __uint128_t m = (__uint128_t)x * n;
        uint64_t l = (uint64_t)m;
        if (l < n) {
            const uint64_t t = (0 - n) % n;
            while (l < t) { x = next(); m = (__uint128_t)x * n; l = (uint64_t)m; }
        }
        return (uint64_t)(m >> 64);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

struct EdgeList {
    std::vector<int64_t> src, dst;
};

void finish(std::vector<std::pair<int64_t, int64_t>>& ed, EdgeList& out) {
    for (auto& e : ed) if (e.first > e.second) std::swap(e.first, e.second);
    std::sort(ed.begin(), ed.end());
    ed.erase(std::unique(ed.begin(), ed.end()), ed.end());
    out.src.resize(ed.size()); out.dst.resize(ed.size());
    for (size_t i = 0; i < ed.size(); ++i) { out.src[i] = ed[i].first + 1; out.dst[i] = ed[i].second + 1; }
}

}  // namespace

struct nd_edgelist { EdgeList el; };

extern "C" {

// G(n, m): m distinct unordered pairs, uniformly.
int nd_graph_erdos_renyi(int64_t n, int64_t m, uint64_t seed, nd_edgelist** out) {
    if (!out || n < 2 || m < 0 || m > n * (n - 1) / 2) return nd::set_error(ND_ERR_ARGUMENT, "erdos_renyi: bad n/m");
    Rng r(seed);
    std::unordered_set<uint64_t> seen;
    seen.reserve((size_t)m * 2);
    std::vector<std::pair<int64_t, int64_t>> ed;
    ed.reserve((size_t)m);
    while ((int64_t)ed.size() < m) {
        int64_t a = (int64_t)r.below((uint64_t)n), b = (int64_t)r.below((uint64_t)n);
        if (a == b) continue;
        if (a > b) std::swap(a, b);
        if (seen.insert((uint64_t)a * (uint64_t)n + (uint64_t)b).second) ed.emplace_back(a, b);
    }
    nd_edgelist* h = new (std::nothrow) nd_edgelist;
    if (!h) return nd::set_error(ND_ERR_ALLOC, "allocation failed");
    finish(ed, h->el);
    *out = h;
    return ND_OK;
}

// Barabasi-Albert growth: m initial isolated vertices; every later vertex attaches to
// m distinct existing vertices drawn with probability proportional to degree (the
// first new vertex attaches to all m initial ones).  ne = m*(n-m).
int nd_graph_barabasi_albert(int64_t n, int64_t m, uint64_t seed, nd_edgelist** out) {
    if (!out || m < 1 || n <= m) return nd::set_error(ND_ERR_ARGUMENT, "barabasi_albert: need n > m >= 1");
    Rng r(seed);
    std::vector<int64_t> ends;            // each vertex repeated degree times
    ends.reserve((size_t)(2 * m * (n - m)));
    std::vector<std::pair<int64_t, int64_t>> ed;
    ed.reserve((size_t)(m * (n - m)));
    std::vector<int64_t> pick((size_t)m);
    std::vector<int64_t> mark((size_t)n, -1);
    for (int64_t v = m; v < n; ++v) {
        if (v == m) {
            for (int64_t j = 0; j < m; ++j) pick[(size_t)j] = j;
        } else {
            int64_t got = 0;
            while (got < m) {
                const int64_t t = ends[(size_t)r.below(ends.size())];
                if (mark[(size_t)t] == v) continue;
                mark[(size_t)t] = v;
                pick[(size_t)got++] = t;
            }
        }
        for (int64_t j = 0; j < m; ++j) {
            ed.emplace_back(pick[(size_t)j], v);
            ends.push_back(pick[(size_t)j]);
            ends.push_back(v);
        }
    }
    nd_edgelist* h = new (std::nothrow) nd_edgelist;
    if (!h) return nd::set_error(ND_ERR_ALLOC, "allocation failed");
    finish(ed, h->el);
    *out = h;
    return ND_OK;
}

// Random geometric graph: n uniform points in the unit square, an edge for every
// pair closer than radius.  Point coordinates are returned if xy != NULL (2n doubles).
int nd_graph_random_geometric(int64_t n, double radius, uint64_t seed, double* xy, nd_edgelist** out) {
    if (!out || n < 2 || !(radius > 0.0)) return nd::set_error(ND_ERR_ARGUMENT, "random_geometric: bad n/radius");
    Rng r(seed);
    std::vector<double> x((size_t)n), y((size_t)n);
    for (int64_t i = 0; i < n; ++i) { x[(size_t)i] = r.uniform(); y[(size_t)i] = r.uniform(); }
    if (xy) for (int64_t i = 0; i < n; ++i) { xy[2 * i] = x[(size_t)i]; xy[2 * i + 1] = y[(size_t)i]; }
    const int64_t nc = std::max<int64_t>(1, std::min<int64_t>((int64_t)std::floor(1.0 / radius), 1 << 15));
    std::vector<int64_t> head((size_t)(nc * nc), -1), next((size_t)n, -1);
    auto cell = [&](double c) { int64_t k = (int64_t)(c * (double)nc); return k >= nc ? nc - 1 : k; };
    for (int64_t i = n - 1; i >= 0; --i) {
        const int64_t c = cell(x[(size_t)i]) * nc + cell(y[(size_t)i]);
        next[(size_t)i] = head[(size_t)c];
        head[(size_t)c] = i;
    }
    const double r2 = radius * radius;
    std::vector<std::pair<int64_t, int64_t>> ed;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t cx = cell(x[(size_t)i]), cy = cell(y[(size_t)i]);
        for (int64_t dx = -1; dx <= 1; ++dx)
            for (int64_t dy = -1; dy <= 1; ++dy) {
                const int64_t ax = cx + dx, ay = cy + dy;
                if (ax < 0 || ay < 0 || ax >= nc || ay >= nc) continue;
                for (int64_t j = head[(size_t)(ax * nc + ay)]; j >= 0; j = next[(size_t)j]) {
                    if (j <= i) continue;
                    const double ddx = x[(size_t)i] - x[(size_t)j], ddy = y[(size_t)i] - y[(size_t)j];
                    if (ddx * ddx + ddy * ddy < r2) ed.emplace_back(i, j);
                }
            }
    }
    nd_edgelist* h = new (std::nothrow) nd_edgelist;
    if (!h) return nd::set_error(ND_ERR_ALLOC, "allocation failed");
    finish(ed, h->el);
    *out = h;
    return ND_OK;
}

// Watts-Strogatz: ring lattice with k nearest neighbours (k even: k/2 per side),
// each lattice edge (i, i+j) rewired with probability beta to (i, uniform w)
// avoiding self-loops and duplicates.
int nd_graph_watts_strogatz(int64_t n, int64_t k, double beta, uint64_t seed, nd_edgelist** out) {
    if (!out || n < 3 || k < 2 || k >= n || (k % 2)) return nd::set_error(ND_ERR_ARGUMENT, "watts_strogatz: need even 2 <= k < n");
    Rng r(seed);
    std::unordered_set<uint64_t> seen;
    auto key = [n](int64_t a, int64_t b) { if (a > b) std::swap(a, b); return (uint64_t)a * (uint64_t)n + (uint64_t)b; };
    std::vector<std::pair<int64_t, int64_t>> ed;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 1; j <= k / 2; ++j) { ed.emplace_back(i, (i + j) % n); seen.insert(key(i, (i + j) % n)); }
    for (auto& e : ed) {
        if (r.uniform() >= beta) continue;
        for (int tries = 0; tries < 64; ++tries) {
            const int64_t w = (int64_t)r.below((uint64_t)n);
            if (w == e.first || seen.count(key(e.first, w))) continue;
            seen.erase(key(e.first, e.second));
            seen.insert(key(e.first, w));
            e.second = w;
            break;
        }
    }
    nd_edgelist* h = new (std::nothrow) nd_edgelist;
    if (!h) return nd::set_error(ND_ERR_ALLOC, "allocation failed");
    finish(ed, h->el);
    *out = h;
    return ND_OK;
}

int64_t nd_edgelist_size(const nd_edgelist* el) { return el ? (int64_t)el->el.src.size() : 0; }

int nd_edgelist_copy(const nd_edgelist* el, int64_t* src, int64_t* dst) {
    if (!el || !src || !dst) return nd::set_error(ND_ERR_ARGUMENT, "NULL argument");
    std::copy(el->el.src.begin(), el->el.src.end(), src);
    std::copy(el->el.dst.begin(), el->el.dst.end(), dst);
    return ND_OK;
}

void nd_edgelist_free(nd_edgelist* el) { delete el; }

}  // extern "C"
