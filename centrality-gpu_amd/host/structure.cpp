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
// (Codes/bicc.cpp:81) overflows the stack at a million vertices.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "host_graph.h"

namespace bcx {

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

int build_adjacency(bcx_host *h) {
    const int32_t n = h->n;
    const int64_t m = h->m;
    h->adj_ptr.assign((size_t)n + 1, 0);
    for (int64_t e = 0; e < m; ++e) {
        int32_t u = h->uv[2 * e], v = h->uv[2 * e + 1];
        if (u < 0 || u >= n || v < 0 || v >= n) {
            h->err = "edge " + std::to_string(e) + " has a vertex id outside [0, n)";
            return BCX_E_IO;
        }
        h->adj_ptr[u + 1]++;
        h->adj_ptr[v + 1]++;
    }
    for (int32_t i = 0; i < n; ++i) h->adj_ptr[i + 1] += h->adj_ptr[i];
    h->adj.assign((size_t)(2 * m), 0);
    std::vector<int64_t> fill(h->adj_ptr.begin(), h->adj_ptr.end() - 1);
    // Codes/graph.cpp:52-57: adj[u] gets v, then adj[v] gets u, in file order.
    for (int64_t e = 0; e < m; ++e) {
        int32_t u = h->uv[2 * e], v = h->uv[2 * e + 1];
        h->adj[fill[u]++] = v;
        h->adj[fill[v]++] = u;
    }
    return BCX_OK;
}

static int check_simple_connected(bcx_host *h) {
    const int32_t n = h->n;
    for (int64_t e = 0; e < h->m; ++e)
        if (h->uv[2 * e] == h->uv[2 * e + 1]) {
            h->err = "self loop on vertex " + std::to_string(h->uv[2 * e]);
            return BCX_E_GRAPH;
        }
    std::vector<int32_t> mark((size_t)n, -1);
    for (int32_t u = 0; u < n; ++u)
        for (int64_t j = h->adj_ptr[u]; j < h->adj_ptr[u + 1]; ++j) {
            int32_t v = h->adj[j];
            if (mark[v] == u) {
                h->err = "duplicate edge " + std::to_string(u) + " " + std::to_string(v);
                return BCX_E_GRAPH;
            }
            mark[v] = u;
        }
    if (n == 0) return BCX_OK;
    std::vector<int32_t> q;
    q.reserve(n);
    std::vector<uint8_t> seen((size_t)n, 0);
    seen[0] = 1;
    q.push_back(0);
    for (size_t i = 0; i < q.size(); ++i) {
        int32_t u = q[i];
        for (int64_t j = h->adj_ptr[u]; j < h->adj_ptr[u + 1]; ++j)
            if (!seen[h->adj[j]]) {
                seen[h->adj[j]] = 1;
                q.push_back(h->adj[j]);
            }
    }
    if ((int32_t)q.size() != n) {
        h->err = "graph is disconnected (" + std::to_string(q.size()) + " of " +
                 std::to_string(n) + " vertices reachable from 0)";
        return BCX_E_GRAPH;
    }
    return BCX_OK;
}

// Edge-stack Tarjan in the reference's visit order: roots 0..n-1, neighbours
// in file order, tree edge pushed before descending, back edge pushed when
// dfs[v] < dfs[u] and v is not the DFS parent, a component popped when
// dfs[u] <= low[v] after returning from v (Codes/bicc.cpp:81-161).
static void tarjan(bcx_host *h) {
    const int32_t n = h->n;
    std::vector<int32_t> dfs((size_t)n, -1), low((size_t)n, -1);
    std::vector<int64_t> it((size_t)n, 0);
    std::vector<int32_t> parent((size_t)n, -1);
    std::vector<int32_t> stk;            // DFS vertex stack
    std::vector<int32_t> estk;           // edge stack (u,v pairs)
    int32_t last = 0;
    int32_t children_of_0 = 0;           // Codes/bicc.cpp:13,86-87
    h->is_art.assign((size_t)n, 0);
    h->bcc_ptr.assign(1, 0);
    h->bcc_uv.clear();
    h->bcc_nv.clear();
    std::vector<int32_t> vmark((size_t)n, -1);

    for (int32_t r = 0; r < n; ++r) {
        if (dfs[r] != -1) continue;
        dfs[r] = low[r] = ++last;
        parent[r] = -1;
        it[r] = h->adj_ptr[r];
        stk.push_back(r);
        while (!stk.empty()) {
            int32_t u = stk.back();
            if (it[u] < h->adj_ptr[u + 1]) {
                int32_t v = h->adj[it[u]++];
                if (dfs[v] == -1) {
                    estk.push_back(u);
                    estk.push_back(v);
                    dfs[v] = low[v] = ++last;
                    parent[v] = u;
                    if (u == 0) children_of_0++;
                    it[v] = h->adj_ptr[v];
                    stk.push_back(v);
                } else if (dfs[v] < dfs[u] && v != parent[u]) {
                    estk.push_back(u);
                    estk.push_back(v);
                    if (dfs[v] < low[u]) low[u] = dfs[v];
                }
                continue;
            }
            // u finished: return to its parent p with child u
            stk.pop_back();
            int32_t p = parent[u];
            if (p < 0) continue;
            int32_t c = u;
            if (low[c] < low[p]) low[p] = low[c];
            if (dfs[p] <= low[c] && !estk.empty()) {
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
    }
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
            int32_t v = -1;
            (void)v;
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

int decompose(bcx_host *h, bool use_chains, bool check_input) {
    if (h->adj_ptr.size() != (size_t)h->n + 1) {
        int rc = build_adjacency(h);
        if (rc) return rc;
    }
    if (check_input) {
        int rc = check_simple_connected(h);
        if (rc) return rc;
    }
    tarjan(h);
    std::vector<std::vector<std::pair<int32_t, int32_t>>> reach_lists;
    block_cut(h, reach_lists);

    const int32_t nb = (int32_t)h->bcc_nv.size();
    h->comps.clear();
    std::vector<int32_t> local((size_t)h->n, -1);
    for (int32_t b = 0; b < nb; ++b) {
        if (h->bcc_nv[b] <= 2) continue;  // Codes/to_eargraph_construct.cpp:59-60
        Component c;
        c.bcc = b;
        const int64_t e0 = h->bcc_ptr[b], e1 = h->bcc_ptr[b + 1];
        const int64_t ne = e1 - e0;
        // first-appearance local ids (Codes/modified_apsp.cpp:336-357)
        for (int64_t e = e0; e < e1; ++e)
            for (int k = 0; k < 2; ++k) {
                int32_t v = h->bcc_uv[2 * e + k];
                if (local[v] < 0) { local[v] = (int32_t)c.gid.size(); c.gid.push_back(v); }
            }
        c.n = (int32_t)c.gid.size();
        // arcs (u->v),(v->u) per edge, stable counting sort by source
        c.row_ptr.assign((size_t)c.n + 1, 0);
        for (int64_t e = e0; e < e1; ++e) {
            c.row_ptr[local[h->bcc_uv[2 * e]] + 1]++;
            c.row_ptr[local[h->bcc_uv[2 * e + 1]] + 1]++;
        }
        for (int32_t i = 0; i < c.n; ++i) c.row_ptr[i + 1] += c.row_ptr[i];
        c.col.assign((size_t)(2 * ne), 0);
        std::vector<int32_t> fill(c.row_ptr.begin(), c.row_ptr.end() - 1);
        for (int64_t e = e0; e < e1; ++e) {
            int32_t u = local[h->bcc_uv[2 * e]], v = local[h->bcc_uv[2 * e + 1]];
            c.col[fill[u]++] = v;
            c.col[fill[v]++] = u;
        }
        c.reach.assign((size_t)c.n, 0);
        for (auto &pr : reach_lists[b]) c.reach[local[pr.first]] = pr.second;
        for (int32_t v : c.gid) local[v] = -1;
        if (use_chains) extract_chains(c);
        else {
            c.chain_ptr.assign(1, 0);
            c.n_free = c.n;
        }
        h->comps.push_back(std::move(c));
    }
    h->decomposed = true;
    return BCX_OK;
}

}  // namespace bcx

// ------------------------------------------------------------------------
// C-ABI: host half
// ------------------------------------------------------------------------
using bcx::Component;

extern "C" {

int bcx_host_from_edges(int32_t n, int64_t m, const int32_t *uv, bcx_host **out) {
    if (!out || n < 0 || m < 0 || (m > 0 && !uv)) return BCX_E_ARG;
    *out = nullptr;
    bcx_host *h = new bcx_host();
    h->n = n;
    h->m = m;
    h->uv.assign(uv, uv + 2 * m);
    int rc = bcx::build_adjacency(h);
    if (rc) { delete h; return rc; }
    *out = h;
    return BCX_OK;
}

// Text format of Codes/ReadMe:6-13 / Codes/graph.cpp:44-59.  Parsed with a
// hand-rolled scanner (fscanf per token is the reference's bottleneck at
// tens of millions of lines).
int bcx_host_from_file(const char *path, bcx_host **out) {
    if (!path || !out) return BCX_E_ARG;
    *out = nullptr;
    FILE *fp = fopen(path, "rb");
    if (!fp) return BCX_E_IO;
    std::vector<char> buf;
    {
        char tmp[1 << 16];
        size_t k;
        while ((k = fread(tmp, 1, sizeof(tmp), fp)) > 0) buf.insert(buf.end(), tmp, tmp + k);
    }
    fclose(fp);
    buf.push_back('\0');
    const char *p = buf.data();
    auto next_int = [&](long long &x) -> bool {
        while (*p && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
        if (!*p) return false;
        bool neg = false;
        if (*p == '-' || *p == '+') { neg = (*p == '-'); ++p; }
        if (*p < '0' || *p > '9') return false;
        long long v = 0;
        while (*p >= '0' && *p <= '9') { v = v * 10 + (*p - '0'); ++p; }
        x = neg ? -v : v;
        return true;
    };
    long long n, m;
    if (!next_int(n) || !next_int(m) || n < 0 || m < 0 || n > INT32_MAX) return BCX_E_IO;
    std::vector<int32_t> uv;
    uv.reserve((size_t)(2 * m));
    for (long long e = 0; e < m; ++e) {
        long long u, v, w;
        if (!next_int(u) || !next_int(v) || !next_int(w)) return BCX_E_IO;
        if (u < 0 || u >= n || v < 0 || v >= n) return BCX_E_IO;
        if (w != 1) return BCX_E_IO;  // unit weights only (Codes/ReadMe:6-9)
        uv.push_back((int32_t)u);
        uv.push_back((int32_t)v);
    }
    return bcx_host_from_edges((int32_t)n, (int64_t)m, uv.data(), out);
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

void bcx_free(void *p) { free(p); }

}  // extern "C"
