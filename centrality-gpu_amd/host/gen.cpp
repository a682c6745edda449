// Seeded synthetic workloads for the five benchmark graphs.  All outputs are
// undirected, unit weight, connected, free of self loops and duplicate
// edges, with vertex ids shuffled and the edge list in shuffled order.
// RNG: xoshiro256** seeded through SplitMix64.
//
//  1  16 Erdos-Renyi blocks of 128 vertices (p = 8/127, resampled until
//     biconnected) joined in a random tree by shared cut vertices; 20% of
//     edges subdivided into chains of U{1..8} vertices; 512 pendant vertices
//     as random recursive trees.                                  (~9.9k V)
//  2  Chung-Lu core, exponent 2.1 (w_i ~ (i+10)^(-1/1.1), clipped at
//     sqrt(sum w)), 550k vertices and 6M distinct edges, giant component
//     kept; 300k chain vertices (geometric lengths, mean 6) subdividing
//     random core edges; pendant-tree vertices up to exactly 1,000,000 V.
//  3  Largest BCC of a Chung-Lu sample (~500k V / 4M E) plus 40,000
//     Erdos-Renyi BCCs of log-uniform size in [3, 2000] (Hamiltonian cycle
//     plus edges for an average degree U[4,12]) joined by cut vertices.
//  4  256x256 4-neighbour grid, 5% of edges deleted outside a random
//     spanning tree (so connectivity holds), 2,000 chains of U{1..8}
//     vertices spliced into random edges.                         (~74.5k V)
//  5  R-MAT scale 20 (a=.57, b=c=.19, d=.05), edge factor 16, symmetrised,
//     deduplicated, self loops dropped, largest component kept; 10% of 2^20
//     extra vertices as chains of U{2..10} subdividing random edges.
//
// scale (0,1] shrinks vertex and edge budgets for tests.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "host_graph.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9e3779b97f4a7c15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
            s[i] = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }              // small bias is irrelevant here
    int64_t range(int64_t lo, int64_t hi) { return lo + (int64_t)below((uint64_t)(hi - lo + 1)); }
};

using Edges = std::vector<std::pair<int32_t, int32_t>>;

struct UF {
    std::vector<int32_t> p;
    explicit UF(int32_t n) : p((size_t)n) { std::iota(p.begin(), p.end(), 0); }
    int32_t f(int32_t x) {
        while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
        return x;
    }
    bool u(int32_t a, int32_t b) {
        a = f(a); b = f(b);
        if (a == b) return false;
        p[a] = b;
        return true;
    }
};

uint64_t key(int32_t a, int32_t b) {
    if (a > b) std::swap(a, b);
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

// Keeps the largest connected component; relabels to 0..k-1.
void largest_cc(int32_t &n, Edges &E) {
    UF uf(n);
    for (auto &e : E) uf.u(e.first, e.second);
    std::vector<int32_t> cnt((size_t)n, 0);
    for (int32_t v = 0; v < n; ++v) cnt[uf.f(v)]++;
    int32_t best = (int32_t)(std::max_element(cnt.begin(), cnt.end()) - cnt.begin());
    std::vector<int32_t> id((size_t)n, -1);
    int32_t k = 0;
    for (int32_t v = 0; v < n; ++v) if (uf.f(v) == best) id[v] = k++;
    Edges out;
    out.reserve(E.size());
    for (auto &e : E) if (id[e.first] >= 0) out.push_back({id[e.first], id[e.second]});
    n = k;
    E.swap(out);
}

// Replaces edge i by a path through `len` new vertices.
void subdivide(int32_t &n, Edges &E, size_t i, int32_t len) {
    int32_t u = E[i].first, v = E[i].second;
    int32_t prev = u;
    for (int32_t k = 0; k < len; ++k) {
        int32_t x = n++;
        if (k == 0) E[i] = {prev, x};
        else E.push_back({prev, x});
        prev = x;
    }
    E.push_back({prev, v});
}

void add_pendant_trees(int32_t &n, Edges &E, int64_t count, Rng &r) {
    for (int64_t k = 0; k < count; ++k) {
        int32_t attach = (int32_t)r.below((uint64_t)n);
        E.push_back({attach, n++});
    }
}

// Distinct edges drawn with endpoint probabilities proportional to w.
Edges chung_lu(int32_t N, int64_t E_target, double i0, Rng &r) {
    std::vector<double> w((size_t)N);
    double sum = 0;
    for (int32_t i = 0; i < N; ++i) { w[i] = std::pow(i + i0, -1.0 / 1.1); sum += w[i]; }
    const double scale = 2.0 * (double)E_target / sum;
    sum = 0;
    for (auto &x : w) { x *= scale; sum += x; }
    const double cap = std::sqrt(sum);
    sum = 0;
    for (auto &x : w) { x = std::min(x, cap); sum += x; }
    // alias table
    std::vector<double> prob((size_t)N);
    std::vector<int32_t> alias((size_t)N);
    std::vector<int32_t> small, large;
    for (int32_t i = 0; i < N; ++i) {
        prob[i] = w[i] * N / sum;
        (prob[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        int32_t s = small.back(); small.pop_back();
        int32_t l = large.back();
        alias[s] = l;
        prob[l] -= 1.0 - prob[s];
        if (prob[l] < 1.0) { large.pop_back(); small.push_back(l); }
    }
    for (int32_t i : small) prob[i] = 1.0;
    for (int32_t i : large) prob[i] = 1.0;
    auto draw = [&]() -> int32_t {
        int32_t i = (int32_t)r.below((uint64_t)N);
        return r.uni() < prob[i] ? i : alias[i];
    };
    std::vector<uint64_t> keys;
    keys.reserve((size_t)(E_target * 1.2));
    while ((int64_t)keys.size() < E_target) {
        int64_t want = (E_target - (int64_t)keys.size()) * 11 / 10 + 16;
        for (int64_t k = 0; k < want; ++k) {
            int32_t a = draw(), b = draw();
            if (a != b) keys.push_back(key(a, b));
        }
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    }
    // keep a random subset of exactly E_target
    for (size_t i = keys.size(); i > 1; --i) std::swap(keys[i - 1], keys[r.below(i)]);
    keys.resize((size_t)E_target);
    Edges out;
    out.reserve(keys.size());
    for (uint64_t k : keys) out.push_back({(int32_t)(k >> 32), (int32_t)(k & 0xffffffffu)});
    return out;
}

bool biconnected(int32_t n, const Edges &E) {
    bcx_host *h = nullptr;
    std::vector<int32_t> uv;
    for (auto &e : E) { uv.push_back(e.first); uv.push_back(e.second); }
    if (bcx_host_from_edges(n, (int64_t)E.size(), uv.data(), &h) != BCX_OK) return false;
    bool ok = bcx_host_decompose(h, 0, 1) == BCX_OK && bcx_host_num_bcc(h) == 1;
    bcx_host_free(h);
    return ok;
}

// Largest BCC of (n, E), relabelled.
void largest_bcc(int32_t &n, Edges &E) {
    bcx_host *h = nullptr;
    std::vector<int32_t> uv;
    for (auto &e : E) { uv.push_back(e.first); uv.push_back(e.second); }
    bcx_host_from_edges(n, (int64_t)E.size(), uv.data(), &h);
    bcx_host_decompose(h, 0, 0);
    int32_t nb = bcx_host_num_bcc(h), best = 0;
    int64_t best_e = -1;
    for (int32_t b = 0; b < nb; ++b) {
        int64_t e;
        bcx_host_bcc_size(h, b, nullptr, &e);
        if (e > best_e) { best_e = e; best = b; }
    }
    std::vector<int32_t> buv((size_t)(2 * best_e));
    bcx_host_bcc_edges(h, best, buv.data());
    bcx_host_free(h);
    std::vector<int32_t> id((size_t)n, -1);
    int32_t k = 0;
    Edges out;
    for (int64_t e = 0; e < best_e; ++e) {
        int32_t a = buv[2 * e], b = buv[2 * e + 1];
        if (id[a] < 0) id[a] = k++;
        if (id[b] < 0) id[b] = k++;
        out.push_back({id[a], id[b]});
    }
    n = k;
    E.swap(out);
}

Edges er_block(int32_t s, double avg_deg, Rng &r, bool ham) {
    Edges E;
    if (ham || s == 3) {
        std::vector<int32_t> perm((size_t)s);
        std::iota(perm.begin(), perm.end(), 0);
        for (int32_t i = s; i > 1; --i) std::swap(perm[i - 1], perm[r.below((uint64_t)i)]);
        for (int32_t i = 0; i < s; ++i) E.push_back({perm[i], perm[(i + 1) % s]});
        if (s == 3) return E;
        double p = std::max(0.0, (avg_deg - 2.0) / (double)(s - 1));
        std::vector<uint64_t> have;
        for (auto &e : E) have.push_back(key(e.first, e.second));
        std::sort(have.begin(), have.end());
        // geometric skipping over the s(s-1)/2 pairs
        if (p > 0) {
            int64_t total = (int64_t)s * (s - 1) / 2, idx = -1;
            const double lq = std::log(1.0 - std::min(p, 0.999999));
            while (true) {
                double u = r.uni();
                idx += 1 + (int64_t)std::floor(std::log(1.0 - u) / lq);
                if (idx >= total) break;
                // pair index -> (a, b), a < b
                int64_t a = (int64_t)((2.0 * s - 1 - std::sqrt((2.0 * s - 1) * (2.0 * s - 1) - 8.0 * idx)) / 2);
                while (a > 0 && a * (2 * s - a - 1) / 2 > idx) --a;
                while ((a + 1) * (2 * s - a - 2) / 2 <= idx) ++a;
                int64_t b = idx - a * (2 * s - a - 1) / 2 + a + 1;
                if (!std::binary_search(have.begin(), have.end(), key((int32_t)a, (int32_t)b)))
                    E.push_back({(int32_t)a, (int32_t)b});
            }
        }
        return E;
    }
    double p = avg_deg / (double)(s - 1);
    for (int32_t a = 0; a < s; ++a)
        for (int32_t b = a + 1; b < s; ++b)
            if (r.uni() < p) E.push_back({a, b});
    return E;
}

void finish(int32_t n, Edges &E, Rng &r, int32_t *n_out, int64_t *m_out, int32_t **uv_out) {
    std::vector<int32_t> perm((size_t)n);
    std::iota(perm.begin(), perm.end(), 0);
    for (int32_t i = n; i > 1; --i) std::swap(perm[i - 1], perm[r.below((uint64_t)i)]);
    for (size_t i = E.size(); i > 1; --i) std::swap(E[i - 1], E[r.below(i)]);
    int32_t *uv = (int32_t *)malloc(sizeof(int32_t) * 2 * std::max<size_t>(E.size(), 1));
    for (size_t i = 0; i < E.size(); ++i) {
        int32_t a = perm[E[i].first], b = perm[E[i].second];
        if (r.next() & 1) std::swap(a, b);
        uv[2 * i] = a;
        uv[2 * i + 1] = b;
    }
    *n_out = n;
    *m_out = (int64_t)E.size();
    *uv_out = uv;
}

int gen1(Rng &r, double scale, int32_t &n, Edges &E) {
    const int32_t blocks = std::max(2, (int32_t)std::lround(16 * scale));
    const int32_t bs = 128;
    n = 0;
    std::vector<std::vector<int32_t>> members;
    for (int32_t b = 0; b < blocks; ++b) {
        Edges be;
        do { be = er_block(bs, 8.0, r, false); } while (!biconnected(bs, be));
        std::vector<int32_t> id((size_t)bs);
        if (b == 0) {
            for (int32_t i = 0; i < bs; ++i) id[i] = n++;
        } else {
            const auto &pb = members[r.below((uint64_t)b)];
            id[0] = pb[r.below(pb.size())];
            for (int32_t i = 1; i < bs; ++i) id[i] = n++;
        }
        members.push_back(id);
        for (auto &e : be) E.push_back({id[e.first], id[e.second]});
    }
    const size_t m0 = E.size();
    std::vector<size_t> idx(m0);
    std::iota(idx.begin(), idx.end(), 0);
    for (size_t i = m0; i > 1; --i) std::swap(idx[i - 1], idx[r.below(i)]);
    for (size_t k = 0; k < m0 / 5; ++k) subdivide(n, E, idx[k], (int32_t)r.range(1, 8));
    add_pendant_trees(n, E, std::max<int64_t>(1, std::lround(512 * scale)), r);
    return BCX_OK;
}

int gen2(Rng &r, double scale, int32_t &n, Edges &E) {
    const int32_t total = (int32_t)std::lround(1000000 * scale);
    n = (int32_t)std::lround(550000 * scale);
    E = chung_lu(n, (int64_t)std::llround(6000000 * scale), 10.0, r);
    largest_cc(n, E);
    int64_t chain_budget = std::llround(300000 * scale);
    const double p = 1.0 / 6.0;
    const size_t m0 = E.size();
    std::vector<size_t> idx(m0);
    std::iota(idx.begin(), idx.end(), 0);
    for (size_t i = m0; i > 1; --i) std::swap(idx[i - 1], idx[r.below(i)]);
    size_t k = 0;
    while (chain_budget > 0 && k < m0) {
        int32_t len = 1 + (int32_t)std::floor(std::log(1.0 - r.uni()) / std::log(1.0 - p));
        len = (int32_t)std::min<int64_t>(len, chain_budget);
        subdivide(n, E, idx[k++], len);
        chain_budget -= len;
    }
    add_pendant_trees(n, E, std::max<int64_t>(0, (int64_t)total - n), r);
    return BCX_OK;
}

int gen3(Rng &r, double scale, int32_t &n, Edges &E) {
    n = (int32_t)std::lround(560000 * scale);
    E = chung_lu(n, (int64_t)std::llround(4500000 * scale), 10.0, r);
    largest_bcc(n, E);
    std::vector<std::vector<int32_t>> members;
    {
        // members of block 0: sample up to 4096 vertices for attachment
        std::vector<int32_t> m0;
        for (int32_t i = 0; i < n; ++i) m0.push_back(i);
        members.push_back(m0);
    }
    const int32_t nblocks = std::max(1, (int32_t)std::lround(40000 * scale));
    for (int32_t b = 0; b < nblocks; ++b) {
        int32_t s = (int32_t)std::lround(std::exp(std::log(3.0) + r.uni() * (std::log(2000.0) - std::log(3.0))));
        s = std::max(3, std::min(2000, s));
        double deg = 4.0 + 8.0 * r.uni();
        deg = std::min(deg, (double)(s - 1));
        Edges be = er_block(s, deg, r, true);
        std::vector<int32_t> id((size_t)s);
        const auto &pb = members[r.below(members.size())];
        id[0] = pb[r.below(pb.size())];
        for (int32_t i = 1; i < s; ++i) id[i] = n++;
        for (auto &e : be) E.push_back({id[e.first], id[e.second]});
        members.push_back(std::move(id));
    }
    return BCX_OK;
}

int gen4(Rng &r, double scale, int32_t &n, Edges &E) {
    const int32_t side = std::max(4, (int32_t)std::lround(256 * std::sqrt(scale)));
    n = side * side;
    for (int32_t y = 0; y < side; ++y)
        for (int32_t x = 0; x < side; ++x) {
            int32_t v = y * side + x;
            if (x + 1 < side) E.push_back({v, v + 1});
            if (y + 1 < side) E.push_back({v, v + side});
        }
    // random spanning tree (Kruskal over shuffled edges) is kept intact
    std::vector<size_t> idx(E.size());
    std::iota(idx.begin(), idx.end(), 0);
    for (size_t i = idx.size(); i > 1; --i) std::swap(idx[i - 1], idx[r.below(i)]);
    UF uf(n);
    std::vector<uint8_t> tree(E.size(), 0);
    for (size_t i : idx) if (uf.u(E[i].first, E[i].second)) tree[i] = 1;
    std::vector<size_t> cand;
    for (size_t i = 0; i < E.size(); ++i) if (!tree[i]) cand.push_back(i);
    for (size_t i = cand.size(); i > 1; --i) std::swap(cand[i - 1], cand[r.below(i)]);
    const size_t del = std::min(cand.size(), (size_t)std::llround(0.05 * (double)E.size()));
    std::vector<uint8_t> gone(E.size(), 0);
    for (size_t k = 0; k < del; ++k) gone[cand[k]] = 1;
    Edges keep;
    for (size_t i = 0; i < E.size(); ++i) if (!gone[i]) keep.push_back(E[i]);
    E.swap(keep);
    const int32_t chains = std::max(1, (int32_t)std::lround(2000 * scale));
    std::vector<size_t> sel(E.size());
    std::iota(sel.begin(), sel.end(), 0);
    for (size_t i = sel.size(); i > 1; --i) std::swap(sel[i - 1], sel[r.below(i)]);
    for (int32_t k = 0; k < chains && k < (int32_t)sel.size(); ++k)
        subdivide(n, E, sel[k], (int32_t)r.range(1, 8));
    return BCX_OK;
}

int gen5(Rng &r, double scale, int32_t &n, Edges &E) {
    int sc = 20;
    while (sc > 8 && (double)(1 << sc) > (double)(1 << 20) * scale * 1.0001) --sc;
    n = 1 << sc;
    const int64_t draws = (int64_t)16 * n;
    const double a = 0.57, b = 0.19, c = 0.19;
    std::vector<uint64_t> keys;
    keys.reserve((size_t)draws);
    for (int64_t k = 0; k < draws; ++k) {
        int32_t u = 0, v = 0;
        for (int bit = 0; bit < sc; ++bit) {
            double x = r.uni();
            int ub = 0, vb = 0;
            if (x < a) {}
            else if (x < a + b) vb = 1;
            else if (x < a + b + c) ub = 1;
            else { ub = 1; vb = 1; }
            u |= ub << bit;
            v |= vb << bit;
        }
        if (u != v) keys.push_back(key(u, v));
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    E.clear();
    for (uint64_t k : keys) E.push_back({(int32_t)(k >> 32), (int32_t)(k & 0xffffffffu)});
    largest_cc(n, E);
    int64_t budget = (int64_t)std::llround(0.1 * (double)(1 << sc));
    const size_t m0 = E.size();
    std::vector<size_t> idx(m0);
    std::iota(idx.begin(), idx.end(), 0);
    for (size_t i = m0; i > 1; --i) std::swap(idx[i - 1], idx[r.below(i)]);
    size_t k = 0;
    while (budget > 0 && k < m0) {
        int32_t len = (int32_t)std::min<int64_t>(r.range(2, 10), budget);
        subdivide(n, E, idx[k++], len);
        budget -= len;
    }
    return BCX_OK;
}

}  // namespace

extern "C" int bcx_generate(int32_t config, uint64_t seed, double scale, int32_t *n_out,
                            int64_t *m_out, int32_t **uv_out) {
    if (!n_out || !m_out || !uv_out || !(scale > 0.0) || scale > 1.0) return BCX_E_ARG;
    static const uint64_t seeds[6] = {0, 0x5EED0001ull, 0x5EED0002ull, 0x5EED0003ull, 0x5EED0004ull,
                                      0x5EED0005ull};
    if (config < 1 || config > 5) return BCX_E_ARG;
    Rng r(seed ? seed : seeds[config]);
    int32_t n = 0;
    Edges E;
    int rc;
    switch (config) {
        case 1: rc = gen1(r, scale, n, E); break;
        case 2: rc = gen2(r, scale, n, E); break;
        case 3: rc = gen3(r, scale, n, E); break;
        case 4: rc = gen4(r, scale, n, E); break;
        default: rc = gen5(r, scale, n, E); break;
    }
    if (rc) return rc;
    finish(n, E, r, n_out, m_out, uv_out);
    return BCX_OK;
}
