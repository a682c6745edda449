// Structural reduction of the input graph (the reference's L4 host layer):
//   * edge-list reader                         Codes/graph.cpp:38-72
//   * Tarjan biconnected components           Codes/bicc.cpp:39-161
//   * per-component local relabel             Codes/modified_apsp.cpp:266-363
//   * block-cut tree, reach r_B(v), cut term  Codes/modified_apsp.cpp:77-220
//   * component CSR                           Codes/to_eargraph_construct.cpp:14-42,81-96
//                                             Codes/graph_ear_decompos.h:36-83
//   * degree-2 chains (ear pieces)            Codes/bcc_bfs.cpp:100-236, 928-1311
//
// Everything is iterative and linear in n + m; the reference's recursive DFS
// (Codes/bicc.cpp:81) overflows the stack at a million vertices.  The passes
// that are independent per edge, vertex or component run on host threads
// (BCX_HOST_THREADS, else OMP_NUM_THREADS, else the hardware count); the DFS
// is sequential and keeps the reference's visit order.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "host_graph.h"

namespace bcx {

static int host_threads() {
    for (const char *k : {"BCX_HOST_THREADS", "OMP_NUM_THREADS"}) {
        const char *e = getenv(k);
        if (e && atoi(e) > 0) return std::min(atoi(e), 64);
    }
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hc, 64u));
}

// f(t, begin, end) over T contiguous ranges of [0, n); inline when small.
template <class F>
static void parallel_ranges(int64_t n, int T, F f) {
    if (T <= 1 || n < (1 << 16)) {
        f(0, (int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T);
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] { f(t, n * t / T, n * (t + 1) / T); });
    for (auto &x : th) x.join();
}

// f(t) on T threads (inline when T is 1).
template <class F>
static void parallel_tasks(int T, F f) {
    if (T <= 1) {
        f(0);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T);
    for (int t = 0; t < T; ++t) th.emplace_back([&, t] { f(t); });
    for (auto &x : th) x.join();
}

bcx_component Component::view() const {
    bcx_component c{};
    c.n = n;
    c.arcs = (int64_t)col.size();
    c.row_ptr = row_ptr.data();
    c.col = col.data();
    c.reach = reach.data();
    c.n_chains = (int32_t)(chain_ends.size() / 2);
    c.chain_ptr = chain_ptr.data();
    c.chain_nodes = chain_nodes.data();
    c.chain_ends = chain_ends.data();
    return c;
}

// Adjacency in file order (Codes/graph.cpp:52-57: adj[u] gets v, then adj[v]
// gets u, edge by edge).  The edges are cut into T contiguous ranges; each
// range counts its arcs per vertex, the counts give every (range, vertex) its
// first slot, and each range fills its slots in edge order, so the lists are
// the sequential ones.
int build_adjacency(bcx_host *h) {
    const int32_t n = h->n;
    const int64_t m = h->m;
    const int32_t *uv = h->uv.data();
    // per-range counters cost 4 bytes per (range, vertex): at most ~1 GiB
    int T = host_threads();
    while (T > 1 && (double)T * n * 4.0 > (double)(1ll << 30)) --T;
    if (m < (1 << 16)) T = 1;
    std::atomic<int64_t> bad(INT64_MAX);
    std::vector<std::vector<int32_t>> cnt((size_t)T);
    parallel_ranges(m, T, [&](int t, int64_t e0, int64_t e1) {
        std::vector<int32_t> &c = cnt[t];
        c.assign((size_t)n, 0);
        for (int64_t e = e0; e < e1; ++e) {
            const int32_t u = uv[2 * e], v = uv[2 * e + 1];
            if ((uint32_t)u >= (uint32_t)n || (uint32_t)v >= (uint32_t)n) {
                int64_t cur = bad.load();
                while (e < cur && !bad.compare_exchange_weak(cur, e)) {}
                return;
            }
            c[u]++;
            c[v]++;
        }
    });
    if (bad.load() != INT64_MAX) {
        h->err = "edge " + std::to_string(bad.load()) + " has a vertex id outside [0, n)";
        return BCX_E_IO;
    }
    h->adj_ptr.assign((size_t)n + 1, 0);
    int64_t *ap = h->adj_ptr.data();
    parallel_ranges(n, T, [&](int, int64_t v0, int64_t v1) {
        for (int64_t v = v0; v < v1; ++v) {
            int64_t d = 0;
            for (int t = 0; t < T; ++t) d += cnt[t][v];
            ap[v + 1] = d;
        }
    });
    for (int32_t i = 0; i < n; ++i) ap[i + 1] += ap[i];
    // cnt[t][v] becomes range t's first slot in v's list (fits in int32: it
    // is an offset inside the list)
    parallel_ranges(n, T, [&](int, int64_t v0, int64_t v1) {
        for (int64_t v = v0; v < v1; ++v) {
            int32_t run = 0;
            for (int t = 0; t < T; ++t) {
                const int32_t c = cnt[t][v];
                cnt[t][v] = run;
                run += c;
            }
        }
    });
    h->adj.assign((size_t)(2 * m), 0);
    int32_t *adj = h->adj.data();
    parallel_ranges(m, T, [&](int t, int64_t e0, int64_t e1) {
        std::vector<int32_t> &c = cnt[t];
        for (int64_t e = e0; e < e1; ++e) {
            const int32_t u = uv[2 * e], v = uv[2 * e + 1];
            adj[ap[u] + c[u]++] = v;
            adj[ap[v] + c[v]++] = u;
        }
    });
    return BCX_OK;
}

// Self loops (first by edge order) and duplicate edges (first vertex u whose
// list repeats a neighbour, reported at the repeat).  Connectivity is checked
// from the DFS of vertex 0 in tarjan().
static int check_simple(bcx_host *h) {
    const int32_t n = h->n;
    const int T = host_threads();
    std::atomic<int64_t> loop(INT64_MAX), dup(INT64_MAX);
    auto lower = [](std::atomic<int64_t> &a, int64_t x) {
        int64_t cur = a.load();
        while (x < cur && !a.compare_exchange_weak(cur, x)) {}
    };
    parallel_ranges(h->m, T, [&](int, int64_t e0, int64_t e1) {
        for (int64_t e = e0; e < e1; ++e)
            if (h->uv[2 * e] == h->uv[2 * e + 1]) { lower(loop, e); return; }
    });
    if (loop.load() != INT64_MAX) {
        h->err = "self loop on vertex " + std::to_string(h->uv[2 * loop.load()]);
        return BCX_E_GRAPH;
    }
    const int64_t *ap = h->adj_ptr.data();
    const int32_t *adj = h->adj.data();
    parallel_ranges(n, T, [&](int, int64_t u0, int64_t u1) {
        std::vector<int32_t> tmp;
        for (int64_t u = u0; u < u1 && u < dup.load(std::memory_order_relaxed); ++u) {
            const int64_t d = ap[u + 1] - ap[u];
            if (d < 2) continue;
            tmp.assign(adj + ap[u], adj + ap[u + 1]);
            std::sort(tmp.begin(), tmp.end());
            if (std::adjacent_find(tmp.begin(), tmp.end()) != tmp.end()) { lower(dup, u); return; }
        }
    });
    if (dup.load() != INT64_MAX) {
        const int32_t u = (int32_t)dup.load();
        std::vector<int32_t> seen;
        for (int64_t j = ap[u]; j < ap[u + 1]; ++j) {
            const int32_t v = adj[j];
            if (std::find(seen.begin(), seen.end(), v) != seen.end()) {
                h->err = "duplicate edge " + std::to_string(u) + " " + std::to_string(v);
                return BCX_E_GRAPH;
            }
            seen.push_back(v);
        }
    }
    return BCX_OK;
}

// Edge-stack Tarjan in the reference's visit order: roots 0..n-1, neighbours
// in file order, tree edge pushed before descending, back edge pushed when
// dfs[v] < dfs[u] and v is not the DFS parent, a component popped when
// dfs[u] <= low[v] after returning from v (Codes/bicc.cpp:81-161).
// The per-vertex DFS state is one 32-byte record (one cache line touched per
// visit); the records of the next neighbours in a list and the adjacency of
// the unvisited ones are prefetched: the walk is bound by those random reads.  Returns the number of vertices
// reached from vertex 0.
struct DfsRec {
    int32_t dfs, low, parent, pad;
    int64_t it, end;  // next arc to scan, end of the list
};
static int32_t tarjan(bcx_host *h) {
    const int32_t n = h->n;
    constexpr int PF = 4;  // prefetch distance (arcs) of the neighbour scan
    std::vector<DfsRec> rec((size_t)n);
    DfsRec *R = rec.data();
    const int32_t *adj = h->adj.data();
    for (int32_t v = 0; v < n; ++v) R[v] = DfsRec{-1, -1, -1, 0, h->adj_ptr[v], h->adj_ptr[v + 1]};
    std::vector<int32_t> stk;            // DFS vertex stack
    std::vector<int32_t> estk;           // edge stack (u,v pairs)
    int32_t last = 0, reached0 = 0;
    int32_t children_of_0 = 0;           // Codes/bicc.cpp:13,86-87
    h->is_art.assign((size_t)n, 0);
    h->bcc_ptr.assign(1, 0);
    h->bcc_uv.clear();
    h->bcc_nv.clear();
    std::vector<int32_t> vmark((size_t)n, -1);

    for (int32_t r = 0; r < n; ++r) {
        if (R[r].dfs != -1) continue;
        R[r].dfs = R[r].low = ++last;
        R[r].parent = -1;
        stk.push_back(r);
        for (int64_t j = R[r].it; j < std::min(R[r].end, R[r].it + 2 * PF); ++j) __builtin_prefetch(&R[adj[j]]);
        while (!stk.empty()) {
            const int32_t u = stk.back();
            DfsRec &U = R[u];
            if (U.it < U.end) {
                // records 2*PF arcs ahead; PF ahead, the record is cached, and
                // an unvisited neighbour (a likely descent) gets its list's
                // first line prefetched too
                if (U.it + 2 * PF < U.end) __builtin_prefetch(&R[adj[U.it + 2 * PF]]);
                if (U.it + PF < U.end) {
                    const DfsRec &W = R[adj[U.it + PF]];
                    if (W.dfs == -1) __builtin_prefetch(&adj[W.it]);
                }
                const int32_t v = adj[U.it++];
                DfsRec &V = R[v];
                if (V.dfs == -1) {
                    estk.push_back(u);
                    estk.push_back(v);
                    V.dfs = V.low = ++last;
                    V.parent = u;
                    if (u == 0) children_of_0++;
                    stk.push_back(v);
                    for (int64_t j = V.it; j < std::min(V.end, V.it + 2 * PF); ++j) __builtin_prefetch(&R[adj[j]]);
                } else if (V.dfs < U.dfs && v != U.parent) {
                    estk.push_back(u);
                    estk.push_back(v);
                    if (V.dfs < U.low) U.low = V.dfs;
                }
                continue;
            }
            // u finished: return to its parent p with child u
            stk.pop_back();
            const int32_t p = U.parent;
            if (p < 0) continue;
            const int32_t c = u;
            DfsRec &P = R[p];
            if (U.low < P.low) P.low = U.low;
            if (P.dfs <= U.low && !estk.empty()) {
                int32_t bcc_id = (int32_t)h->bcc_nv.size();
                int32_t nv = 0;
                while (!estk.empty()) {
                    int32_t b = estk.back(); estk.pop_back();
                    int32_t a = estk.back(); estk.pop_back();
                    h->bcc_uv.push_back(a);
                    h->bcc_uv.push_back(b);
                    if (vmark[a] != bcc_id) { vmark[a] = bcc_id; nv++; }
                    if (vmark[b] != bcc_id) { vmark[b] = bcc_id; nv++; }
                    if (a == p && b == c) break;
                }
                h->bcc_nv.push_back(nv);
                h->bcc_ptr.push_back((int64_t)h->bcc_uv.size() / 2);
                if (p != 0 || children_of_0 > 1) h->is_art[p] = 1;
            }
        }
        if (r == 0) reached0 = last;
    }
    return reached0;
}

// Block-cut tree: nodes are BCCs [0, nb) and articulation vertices.  Each
// vertex is counted once (articulation vertices at their own node, the rest
// in their unique BCC); subtree sums S give the reach of an articulation
// vertex a seen from BCC B and the sizes of the components of G - a.
// Reach r_B(a) = S(a) - 1 when a is B's child, n - S(B) - 1 when a is B's
// parent; cut(a) = (n-1)^2 - sum_i |C_i|^2 over the components of G - a.
// This is the closed form behind Codes/modified_apsp.cpp:157-220.
static void block_cut(bcx_host *h, std::vector<std::vector<std::pair<int32_t,int32_t>>> &reach_out) {
    const int32_t n = h->n;
    const int32_t nb = (int32_t)h->bcc_nv.size();
    std::vector<int32_t> art_id((size_t)n, -1);
    int32_t na = 0;
    for (int32_t v = 0; v < n; ++v)
        if (h->is_art[v]) art_id[v] = na++;
    // adjacency block <-> art node; node id: block b -> b, art a -> nb + a
    std::vector<std::vector<int32_t>> tadj((size_t)(nb + na));
    std::vector<int64_t> own((size_t)(nb + na), 0);
    std::vector<int32_t> last_block((size_t)n, -1);
    for (int32_t b = 0; b < nb; ++b) {
        for (int64_t e = h->bcc_ptr[b]; e < h->bcc_ptr[b + 1]; ++e)
            for (int k = 0; k < 2; ++k) {
                int32_t v = h->bcc_uv[2 * e + k];
                if (last_block[v] == b) continue;
                last_block[v] = b;
                if (art_id[v] >= 0) {
                    tadj[b].push_back(nb + art_id[v]);
                    tadj[nb + art_id[v]].push_back(b);
                } else {
                    own[b]++;
                }
            }
    }
    for (int32_t a = 0; a < na; ++a) own[nb + a] = 1;
    // BFS order from block 0, then subtree sums bottom-up.
    const int32_t nt = nb + na;
    std::vector<int32_t> par((size_t)nt, -2), order;
    order.reserve(nt);
    if (nt > 0) {
        par[0] = -1;
        order.push_back(0);
        for (size_t i = 0; i < order.size(); ++i) {
            int32_t x = order[i];
            for (int32_t y : tadj[x])
                if (par[y] == -2) { par[y] = x; order.push_back(y); }
        }
    }
    std::vector<int64_t> S(own);
    for (size_t i = order.size(); i-- > 1;) S[par[order[i]]] += S[order[i]];

    h->cut.assign((size_t)n, 0);
    for (int32_t v = 0; v < n; ++v) {
        if (art_id[v] < 0) continue;
        int32_t x = nb + art_id[v];
        int64_t sumsq = 0;
        for (int32_t y : tadj[x]) {
            int64_t c = (par[y] == x) ? S[y] : (int64_t)n - S[x];
            sumsq += c * c;
        }
        h->cut[v] = (int64_t)(n - 1) * (int64_t)(n - 1) - sumsq;
    }
    // reach per (block, articulation vertex)
    reach_out.assign((size_t)nb, {});
    for (int32_t b = 0; b < nb; ++b)
        for (int32_t y : tadj[b]) {
            int64_t r = (par[y] == b) ? S[y] - 1 : (int64_t)n - S[b] - 1;
            reach_out[b].push_back({y - nb, (int32_t)r});
        }
    // translate art index -> vertex
    std::vector<int32_t> art_vertex((size_t)na);
    for (int32_t v = 0; v < n; ++v)
        if (art_id[v] >= 0) art_vertex[art_id[v]] = v;
    for (auto &lst : reach_out)
        for (auto &pr : lst) pr.first = art_vertex[pr.first];
}

// Maximal paths of component-degree-2 vertices between endpoints of degree
// >= 3.  A component with no vertex of degree >= 3 is a plain cycle and has
// no chains (all its vertices stay sources).
static void extract_chains(Component &c) {
    const int32_t n = c.n;
    c.chain_ptr.assign(1, 0);
    c.chain_nodes.clear();
    c.chain_ends.clear();
    auto deg = [&](int32_t v) { return c.row_ptr[v + 1] - c.row_ptr[v]; };
    bool any3 = false;
    for (int32_t v = 0; v < n; ++v) if (deg(v) >= 3) { any3 = true; break; }
    std::vector<uint8_t> role((size_t)n, 0);  // 1 chain vertex, 2 endpoint
    if (any3) {
        std::vector<uint8_t> used((size_t)n, 0);
        std::vector<int32_t> buf;
        for (int32_t L = 0; L < n; ++L) {
            if (deg(L) < 3) continue;
            for (int32_t j = c.row_ptr[L]; j < c.row_ptr[L + 1]; ++j) {
                int32_t x = c.col[j];
                if (deg(x) != 2 || used[x]) continue;
                buf.clear();
                int32_t prev = L, cur = x;
                while (deg(cur) == 2) {
                    used[cur] = 1;
                    buf.push_back(cur);
                    int32_t a = c.col[c.row_ptr[cur]], b = c.col[c.row_ptr[cur] + 1];
                    int32_t nxt = (a != prev) ? a : b;
                    prev = cur;
                    cur = nxt;
                }
                int32_t R = cur;
                if (R == L) continue;  // cannot happen inside a BCC; keep as sources
                for (int32_t y : buf) { c.chain_nodes.push_back(y); role[y] = 1; }
                c.chain_ptr.push_back((int32_t)c.chain_nodes.size());
                c.chain_ends.push_back(L);
                c.chain_ends.push_back(R);
                role[L] = 2;
                role[R] = 2;
            }
        }
    }
    c.n_free = c.n_endpoints = c.n_chain_vertices = 0;
    for (int32_t v = 0; v < n; ++v) {
        if (role[v] == 0) c.n_free++;
        else if (role[v] == 1) c.n_chain_vertices++;
        else c.n_endpoints++;
    }
}

// One component: first-appearance local ids over the BCC's edge list
// (Codes/modified_apsp.cpp:336-357), arcs (u->v),(v->u) per edge in a stable
// counting sort by source, reach of its articulation vertices, chains.
// `local` is all -1 on entry and on return.
static void build_component(const bcx_host *h, int32_t b, const std::vector<std::pair<int32_t, int32_t>> &reach,
                            bool use_chains, std::vector<int32_t> &local, Component &c) {
    c.bcc = b;
    const int64_t e0 = h->bcc_ptr[b], e1 = h->bcc_ptr[b + 1];
    const int64_t ne = e1 - e0;
    const int32_t *buv = h->bcc_uv.data();
    for (int64_t e = e0; e < e1; ++e)
        for (int k = 0; k < 2; ++k) {
            const int32_t v = buv[2 * e + k];
            if (local[v] < 0) { local[v] = (int32_t)c.gid.size(); c.gid.push_back(v); }
        }
    c.n = (int32_t)c.gid.size();
    c.row_ptr.assign((size_t)c.n + 1, 0);
    for (int64_t e = e0; e < e1; ++e) {
        c.row_ptr[local[buv[2 * e]] + 1]++;
        c.row_ptr[local[buv[2 * e + 1]] + 1]++;
    }
    for (int32_t i = 0; i < c.n; ++i) c.row_ptr[i + 1] += c.row_ptr[i];
    c.col.assign((size_t)(2 * ne), 0);
    std::vector<int32_t> fill(c.row_ptr.begin(), c.row_ptr.end() - 1);
    for (int64_t e = e0; e < e1; ++e) {
        const int32_t u = local[buv[2 * e]], v = local[buv[2 * e + 1]];
        c.col[fill[u]++] = v;
        c.col[fill[v]++] = u;
    }
    c.reach.assign((size_t)c.n, 0);
    for (auto &pr : reach) c.reach[local[pr.first]] = pr.second;
    for (int32_t v : c.gid) local[v] = -1;
    if (use_chains) {
        extract_chains(c);
    } else {
        c.chain_ptr.assign(1, 0);
        c.n_free = c.n;
    }
}

int decompose(bcx_host *h, bool use_chains, bool check_input) {
    if (h->cached) {  // already reduced; the edge list was not stored
        if (use_chains == h->chains) return BCX_OK;
        h->err = "host loaded from a structure cache written with another chains setting";
        return BCX_E_ARG;
    }
    if (h->adj_ptr.size() != (size_t)h->n + 1) {
        int rc = build_adjacency(h);
        if (rc) return rc;
    }
    if (check_input) {
        int rc = check_simple(h);
        if (rc) return rc;
    }
    const int32_t reached0 = tarjan(h);
    if (check_input && h->n > 0 && reached0 != h->n) {
        h->err = "graph is disconnected (" + std::to_string(reached0) + " of " +
                 std::to_string(h->n) + " vertices reachable from 0)";
        return BCX_E_GRAPH;
    }
    std::vector<std::vector<std::pair<int32_t, int32_t>>> reach_lists;
    block_cut(h, reach_lists);

    // components (BCCs of more than two vertices, Codes/to_eargraph_construct.cpp:59-60)
    // are built independently; each thread keeps its own global -> local map
    const int32_t nb = (int32_t)h->bcc_nv.size();
    std::vector<int32_t> blocks;
    for (int32_t b = 0; b < nb; ++b) {
        if (h->bcc_nv[b] <= 2) continue;
        // component CSR offsets are int32 (bcx_component.row_ptr)
        if (2 * (h->bcc_ptr[b + 1] - h->bcc_ptr[b]) > (int64_t)INT32_MAX) {
            h->err = "biconnected component " + std::to_string(b) + " has more than 2^31-1 arcs";
            return BCX_E_GRAPH;
        }
        blocks.push_back(b);
    }
    h->comps.clear();
    h->comps.resize(blocks.size());
    int T = host_threads();
    while (T > 1 && (double)T * h->n * 4.0 > (double)(1ll << 30)) --T;
    if (h->m < (1 << 16)) T = 1;
    T = std::max(1, std::min<int>(T, (int)blocks.size()));
    std::atomic<int64_t> next(0);
    auto worker = [&]() {
        std::vector<int32_t> local((size_t)h->n, -1);
        for (int64_t i; (i = next.fetch_add(1)) < (int64_t)blocks.size();)
            build_component(h, blocks[i], reach_lists[blocks[i]], use_chains, local, h->comps[i]);
    };
    if (T == 1) {
        worker();
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t) th.emplace_back(worker);
        for (auto &x : th) x.join();
    }
    h->decomposed = true;
    h->chains = use_chains;
    return BCX_OK;
}

// Host structure over an edge array (taken over, not copied).
static int host_from_uv(int32_t n, int64_t m, std::vector<int32_t> &&uv, bcx_host **out) {
    *out = nullptr;
    bcx_host *h = new bcx_host();
    h->n = n;
    h->m = m;
    h->uv = std::move(uv);
    int rc = build_adjacency(h);
    if (rc) { delete h; return rc; }
    *out = h;
    return BCX_OK;
}

// Read-only view of a whole file: mapped when it is a regular file, else read.
struct FileView {
    const char *data = nullptr;
    size_t size = 0;
    void *map = nullptr;
    std::vector<char> copy;
    bool open(const char *path) {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat st;
        if (fstat(fd, &st) == 0 && S_ISREG(st.st_mode)) {
            size = (size_t)st.st_size;
            if (size == 0) { ::close(fd); data = ""; return true; }
            map = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (map != MAP_FAILED) {
                ::close(fd);
                data = (const char *)map;
                return true;
            }
            map = nullptr;
        }
        char tmp[1 << 16];
        ssize_t k;
        while ((k = ::read(fd, tmp, sizeof(tmp))) > 0) copy.insert(copy.end(), tmp, tmp + k);
        ::close(fd);
        if (k < 0) return false;
        data = copy.data();
        size = copy.size();
        return true;
    }
    void close() {
        if (map) munmap(map, size);
        map = nullptr;
        copy.clear();
        copy.shrink_to_fit();
    }
    ~FileView() { close(); }
};

}  // namespace bcx

// ------------------------------------------------------------------------
// C-ABI: host half
// ------------------------------------------------------------------------
using bcx::Component;

extern "C" {

int bcx_host_from_edges(int32_t n, int64_t m, const int32_t *uv, bcx_host **out) {
    if (!out || n < 0 || m < 0 || (m > 0 && !uv)) return BCX_E_ARG;
    return bcx::host_from_uv(n, m, std::vector<int32_t>(uv, uv + 2 * m), out);
}

// Text format of Codes/ReadMe:6-13 / Codes/graph.cpp:44-59: `n m`, then m
// lines `u v w` (w = 1), read token by token as fscanf does; whatever follows
// the m-th edge is ignored.  The file is mapped and scanned by T threads over
// ranges cut at whitespace: one pass counts each range's tokens (and where
// the valid input ends), the second parses them straight into the edge array
// (fscanf per token is the reference's bottleneck at tens of millions of
// lines).
int bcx_host_from_file(const char *path, bcx_host **out) {
    if (!path || !out) return BCX_E_ARG;
    *out = nullptr;
    bcx::FileView f;
    if (!f.open(path)) return BCX_E_IO;
    const char *buf = f.data;
    const size_t len = f.size;
    auto ws = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; };
    // one token at p < e (p at a non-space); false if it is not an integer
    auto parse = [](const char *&p, const char *e, long long &x) -> bool {
        bool neg = false;
        if (*p == '-' || *p == '+') {
            neg = (*p == '-');
            if (++p >= e) return false;
        }
        if (*p < '0' || *p > '9') return false;
        long long v = 0;
        int digits = 0;
        while (p < e && *p >= '0' && *p <= '9') {
            if (++digits > 18) return false;
            v = v * 10 + (*p - '0');
            ++p;
        }
        x = neg ? -v : v;
        return true;
    };
    int T = bcx::host_threads();
    if (len < (1u << 20)) T = 1;
    // range t is [cut[t], cut[t+1]); inner cuts sit just after whitespace
    std::vector<size_t> cut((size_t)T + 1);
    for (int t = 0; t <= T; ++t) {
        size_t c = len * (size_t)t / (size_t)T;
        while (c < len && c > 0 && !ws(buf[c - 1])) ++c;
        cut[t] = c;
    }
    // scan range t: every valid token goes to emit(index within the range,
    // value); returns the count and whether the valid input stops inside
    auto scan = [&](int t, auto emit) -> std::pair<size_t, bool> {
        const char *p = buf + cut[t], *e = buf + cut[t + 1];
        size_t k = 0;
        for (;;) {
            while (p < e && ws(*p)) ++p;
            if (p >= e) return {k, false};
            long long x;
            if (!parse(p, e, x)) return {k, true};
            emit(k++, x);
            // a token glued to something else ends the valid input after it
            if (p < e && !ws(*p)) return {k, true};
        }
    };
    std::vector<size_t> cnt((size_t)T, 0);
    std::vector<uint8_t> stopped((size_t)T, 0);
    std::vector<long long> head;  // the first tokens (n, m)
    bcx::parallel_tasks(T, [&](int t) {
        auto r = scan(t, [&](size_t k, long long x) {
            if (t == 0 && k < 2) head.push_back(x);
        });
        cnt[t] = r.first;
        stopped[t] = r.second;
    });
    // the valid tokens are those up to the first range that stopped
    std::vector<size_t> start((size_t)T + 1, 0);
    size_t valid = 0;
    for (int t = 0; t < T; ++t) {
        start[t] = valid;
        valid += cnt[t];
        if (stopped[t]) {
            for (int u = t + 1; u <= T; ++u) start[u] = valid;
            break;
        }
        start[t + 1] = valid;
    }
    if (head.size() < 2) {  // n and m may sit in a later range of a tiny first one
        head.clear();
        for (int t = 0; t < T && head.size() < 2; ++t)
            scan(t, [&](size_t, long long x) { if (head.size() < 2) head.push_back(x); });
        if (head.size() < 2 || valid < 2) return BCX_E_IO;
    }
    const long long n = head[0], m = head[1];
    if (n < 0 || m < 0 || n > INT32_MAX) return BCX_E_IO;
    if ((long double)valid < 2.0L + 3.0L * (long double)m) return BCX_E_IO;
    std::vector<int32_t> uv((size_t)(2 * m));
    std::atomic<int> bad(0);
    const size_t last = 2 + 3 * (size_t)m;  // one past the last token used
    bcx::parallel_tasks(T, [&](int t) {
        if (start[t] >= last || (t > 0 && start[t] == start[t - 1] && stopped[t - 1])) return;
        int32_t *dst = uv.data();
        bool ok = true;
        scan(t, [&](size_t k, long long x) {
            const size_t i = start[t] + k;  // global token index
            if (i < 2 || i >= last) return;
            const size_t e = (i - 2) / 3, f = (i - 2) % 3;
            if (f < 2) {
                if (x < 0 || x >= n) ok = false;
                else dst[2 * e + f] = (int32_t)x;
            } else if (x != 1) {
                ok = false;  // unit weights only (Codes/ReadMe:6-9)
            }
        });
        if (!ok) bad = 1;
    });
    if (bad.load()) return BCX_E_IO;
    f.close();
    return bcx::host_from_uv((int32_t)n, (int64_t)m, std::move(uv), out);
}

void bcx_host_free(bcx_host *h) { delete h; }

const char *bcx_host_last_error(const bcx_host *h) { return h ? h->err.c_str() : "null host"; }

int32_t bcx_host_num_vertices(const bcx_host *h) { return h ? h->n : -1; }
int64_t bcx_host_num_edges(const bcx_host *h) { return h ? h->m : -1; }

int bcx_host_decompose(bcx_host *h, int use_chains, int check_input) {
    if (!h) return BCX_E_ARG;
    return bcx::decompose(h, use_chains != 0, check_input != 0);
}

int32_t bcx_host_num_bcc(const bcx_host *h) {
    return (h && h->decomposed) ? (int32_t)h->bcc_nv.size() : -1;
}

int bcx_host_bcc_size(const bcx_host *h, int32_t i, int32_t *n, int64_t *edges) {
    if (!h || !h->decomposed || i < 0 || i >= (int32_t)h->bcc_nv.size()) return BCX_E_ARG;
    if (n) *n = h->bcc_nv[i];
    if (edges) *edges = h->bcc_ptr[i + 1] - h->bcc_ptr[i];
    return BCX_OK;
}

int bcx_host_bcc_edges(const bcx_host *h, int32_t i, int32_t *uv) {
    if (!h || !h->decomposed || !uv || i < 0 || i >= (int32_t)h->bcc_nv.size()) return BCX_E_ARG;
    if (h->cached) return BCX_E_ARG;  // edge lists are not in the cache
    std::memcpy(uv, h->bcc_uv.data() + 2 * h->bcc_ptr[i],
                sizeof(int32_t) * 2 * (h->bcc_ptr[i + 1] - h->bcc_ptr[i]));
    return BCX_OK;
}

int bcx_host_articulation(const bcx_host *h, uint8_t *flags) {
    if (!h || !h->decomposed || !flags) return BCX_E_ARG;
    std::memcpy(flags, h->is_art.data(), h->is_art.size());
    return BCX_OK;
}

int bcx_host_cut_term(const bcx_host *h, int64_t *cut) {
    if (!h || !h->decomposed || !cut) return BCX_E_ARG;
    std::memcpy(cut, h->cut.data(), sizeof(int64_t) * h->cut.size());
    return BCX_OK;
}

int32_t bcx_host_num_components(const bcx_host *h) {
    return (h && h->decomposed) ? (int32_t)h->comps.size() : -1;
}

int bcx_host_component(const bcx_host *h, int32_t c, bcx_component *out,
                       const int32_t **global_ids, int32_t *bcc_index) {
    if (!h || !h->decomposed || c < 0 || c >= (int32_t)h->comps.size()) return BCX_E_ARG;
    const Component &k = h->comps[c];
    if (out) *out = k.view();
    if (global_ids) *global_ids = k.gid.data();
    if (bcc_index) *bcc_index = k.bcc;
    return BCX_OK;
}

int bcx_host_component_counts(const bcx_host *h, int32_t c, int32_t *n_free,
                              int32_t *n_endpoints, int32_t *n_chain_vertices) {
    if (!h || !h->decomposed || c < 0 || c >= (int32_t)h->comps.size()) return BCX_E_ARG;
    const Component &k = h->comps[c];
    if (n_free) *n_free = k.n_free;
    if (n_endpoints) *n_endpoints = k.n_endpoints;
    if (n_chain_vertices) *n_chain_vertices = k.n_chain_vertices;
    return BCX_OK;
}

}  // extern "C"

// ---- structure cache ---------------------------------------------------
// Little-endian file: the header below, then (each array 8-byte aligned)
// bcc_ptr[nb+1] i64, bcc_nv[nb] i32, is_art[n] u8, cut[n] i64, and per
// component a record of counts followed by gid, row_ptr, col, reach,
// chain_ptr, chain_nodes, chain_ends; an end marker closes the file.
namespace {
constexpr char kMagic[8] = {'B', 'C', 'X', 'S', 'T', 'R', 'U', 'C'};
constexpr uint32_t kVersion = 1;
constexpr uint64_t kEnd = 0x444E4558434221ull;  // "!BCXEND"
struct CacheHeader {
    char magic[8];
    uint32_t version, chains;
    int64_t src_size, src_mtime_ns;
    int64_t m;
    int32_t n, nb, ncomp, pad;
};
struct CompRecord {
    int32_t bcc, n, n_free, n_endpoints, n_chain_vertices, n_chains;
    int64_t arcs, n_chain_nodes;
};

bool source_id(const char *path, int64_t &size, int64_t &mtime_ns) {
    size = mtime_ns = 0;
    if (!path) return true;
    struct stat st;
    if (stat(path, &st) != 0) return false;
    size = (int64_t)st.st_size;
    mtime_ns = (int64_t)st.st_mtim.tv_sec * 1000000000ll + st.st_mtim.tv_nsec;
    return true;
}

struct Writer {
    FILE *f;
    int64_t at = 0;
    bool ok = true;
    void raw(const void *p, size_t nbytes) {
        if (nbytes && fwrite(p, 1, nbytes, f) != nbytes) ok = false;
        at += (int64_t)nbytes;
    }
    template <class T> void arr(const T *p, size_t count) {
        raw(p, sizeof(T) * count);
        static const char zero[8] = {0};
        raw(zero, (size_t)((8 - (at & 7)) & 7));
    }
};

struct Reader {
    const char *p, *e;
    bool ok = true;
    template <class T> bool arr(std::vector<T> &out, int64_t count) {
        const int64_t nbytes = (int64_t)sizeof(T) * count;
        const int64_t padded = (nbytes + 7) & ~7ll;
        if (!ok || count < 0 || padded > e - p) return ok = false;
        out.resize((size_t)count);
        if (nbytes) std::memcpy(out.data(), p, (size_t)nbytes);
        p += padded;
        return true;
    }
    template <class T> bool one(T &x) {
        if (!ok || (int64_t)sizeof(T) > e - p) return ok = false;
        std::memcpy(&x, p, sizeof(T));
        p += (sizeof(T) + 7) & ~(size_t)7;
        return true;
    }
};
}  // namespace

extern "C" {

int bcx_host_save(const bcx_host *h, const char *path, const char *source_path) {
    if (!h || !path || !h->decomposed || h->cached) return BCX_E_ARG;
    CacheHeader hd{};
    std::memcpy(hd.magic, kMagic, 8);
    hd.version = kVersion;
    hd.chains = h->chains ? 1 : 0;
    if (!source_id(source_path, hd.src_size, hd.src_mtime_ns)) return BCX_E_IO;
    hd.m = h->m;
    hd.n = h->n;
    hd.nb = (int32_t)h->bcc_nv.size();
    hd.ncomp = (int32_t)h->comps.size();
    // written beside the target and renamed, so a reader never sees a torn file
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return BCX_E_IO;
    std::vector<char> vbuf(1 << 22);
    setvbuf(f, vbuf.data(), _IOFBF, vbuf.size());
    Writer w{f};
    w.arr(&hd, 1);
    w.arr(h->bcc_ptr.data(), h->bcc_ptr.size());
    w.arr(h->bcc_nv.data(), h->bcc_nv.size());
    w.arr(h->is_art.data(), h->is_art.size());
    w.arr(h->cut.data(), h->cut.size());
    for (const Component &c : h->comps) {
        CompRecord r{c.bcc, c.n, c.n_free, c.n_endpoints, c.n_chain_vertices,
                     (int32_t)(c.chain_ends.size() / 2), (int64_t)c.col.size(), (int64_t)c.chain_nodes.size()};
        w.arr(&r, 1);
        w.arr(c.gid.data(), c.gid.size());
        w.arr(c.row_ptr.data(), c.row_ptr.size());
        w.arr(c.col.data(), c.col.size());
        w.arr(c.reach.data(), c.reach.size());
        w.arr(c.chain_ptr.data(), c.chain_ptr.size());
        w.arr(c.chain_nodes.data(), c.chain_nodes.size());
        w.arr(c.chain_ends.data(), c.chain_ends.size());
    }
    w.arr(&kEnd, 1);
    const bool ok = w.ok && fclose(f) == 0;
    if (!ok || rename(tmp.c_str(), path) != 0) {
        remove(tmp.c_str());
        return BCX_E_IO;
    }
    return BCX_OK;
}

int bcx_host_load(const char *path, const char *source_path, int use_chains, bcx_host **out) {
    if (!path || !out) return BCX_E_ARG;
    *out = nullptr;
    bcx::FileView f;
    if (!f.open(path)) return BCX_E_IO;
    Reader r{f.data, f.data + f.size};
    CacheHeader hd;
    if (!r.one(hd) || std::memcmp(hd.magic, kMagic, 8) || hd.version != kVersion) return BCX_E_IO;
    int64_t size, mtime;
    if (!source_id(source_path, size, mtime)) return BCX_E_IO;
    if (source_path && (size != hd.src_size || mtime != hd.src_mtime_ns)) return BCX_E_IO;
    if ((int)hd.chains != (use_chains ? 1 : 0) || hd.n < 0 || hd.nb < 0 || hd.ncomp < 0 || hd.m < 0)
        return BCX_E_IO;
    bcx_host *h = new bcx_host();
    h->n = hd.n;
    h->m = hd.m;
    h->chains = hd.chains != 0;
    h->cached = true;
    auto fail = [&]() { delete h; return BCX_E_IO; };
    if (!r.arr(h->bcc_ptr, (int64_t)hd.nb + 1) || !r.arr(h->bcc_nv, hd.nb) || !r.arr(h->is_art, hd.n) ||
        !r.arr(h->cut, hd.n))
        return fail();
    h->comps.resize((size_t)hd.ncomp);
    for (Component &c : h->comps) {
        CompRecord cr;
        if (!r.one(cr) || cr.n < 0 || cr.n_chains < 0 || cr.arcs < 0 || cr.n_chain_nodes < 0) return fail();
        c.bcc = cr.bcc;
        c.n = cr.n;
        c.n_free = cr.n_free;
        c.n_endpoints = cr.n_endpoints;
        c.n_chain_vertices = cr.n_chain_vertices;
        if (!r.arr(c.gid, cr.n) || !r.arr(c.row_ptr, (int64_t)cr.n + 1) || !r.arr(c.col, cr.arcs) ||
            !r.arr(c.reach, cr.n) || !r.arr(c.chain_ptr, (int64_t)cr.n_chains + 1) ||
            !r.arr(c.chain_nodes, cr.n_chain_nodes) || !r.arr(c.chain_ends, 2 * (int64_t)cr.n_chains))
            return fail();
        // the views handed to the GPU index with these: keep them in range
        if (c.row_ptr[0] != 0 || c.row_ptr[cr.n] != cr.arcs) return fail();
        for (int32_t v = 0; v < cr.n; ++v)
            if (c.row_ptr[v + 1] < c.row_ptr[v] || c.gid[v] < 0 || c.gid[v] >= hd.n) return fail();
        for (int32_t x : c.col)
            if (x < 0 || x >= cr.n) return fail();
        if (c.chain_ptr[0] != 0 || c.chain_ptr[cr.n_chains] != cr.n_chain_nodes) return fail();
        for (int32_t k = 0; k < cr.n_chains; ++k)
            if (c.chain_ptr[k + 1] < c.chain_ptr[k]) return fail();
        for (int32_t x : c.chain_nodes)
            if (x < 0 || x >= cr.n) return fail();
        for (int32_t x : c.chain_ends)
            if (x < 0 || x >= cr.n) return fail();
    }
    uint64_t end = 0;
    if (!r.one(end) || end != kEnd || r.p != r.e) return fail();
    h->decomposed = true;
    *out = h;
    return BCX_OK;
}

void bcx_free(void *p) { free(p); }

}  // extern "C"
