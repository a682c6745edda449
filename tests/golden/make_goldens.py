"""Writes tests/golden/goldens.json: expected outputs computed HERE with
networkx 3.4.2 (run once where networkx is installed; the tests only read
the JSON, so machines without networkx can run them).  Inputs: the reference's own sample inputs
(Codes/sample_input/test1..4.txt, copied beside this script as data) and
small seeded graphs with closed forms or known structure.

Per graph: edge list, 2 x networkx.betweenness_centrality(normalized=False)
(the reference's ordered-pair convention, Codes/modified_apsp.cpp:209),
biconnected components as edge sets, articulation points, and for one source
per graph the BFS levels and shortest-path counts (networkx
single_source_shortest_path_length + all_shortest_paths counts).
"""
import json
import os
import random

import networkx as nx

HERE = os.path.dirname(os.path.abspath(__file__))


def read_sample(path):
    with open(path) as f:
        n, m = map(int, f.readline().split())
        edges = [tuple(map(int, f.readline().split()[:2])) for _ in range(m)]
    return n, edges


def graphs():
    for k in range(1, 5):
        n, e = read_sample(os.path.join(HERE, f"test{k}.txt"))
        yield f"sample_test{k}", n, e
    yield "path7", 7, [(i, i + 1) for i in range(6)]
    yield "star6", 7, [(0, i) for i in range(1, 7)]
    yield "cycle8", 8, [(i, (i + 1) % 8) for i in range(8)]
    yield "k6", 6, [(i, j) for i in range(6) for j in range(i + 1, 6)]
    g = nx.grid_2d_graph(5, 5)
    idx = {v: i for i, v in enumerate(sorted(g.nodes))}
    yield "grid5x5", 25, [(idx[a], idx[b]) for a, b in g.edges]
    g = nx.barbell_graph(5, 3)
    yield "barbell5_3", g.number_of_nodes(), list(g.edges)
    rng = random.Random(20240611)
    for t in range(4):
        while True:
            n = rng.randint(12, 40)
            g = nx.gnp_random_graph(n, rng.uniform(0.08, 0.25), seed=rng.randint(0, 10**9))
            if nx.is_connected(g):
                break
        # subdivide a few edges into chains and hang pendant trees
        e = list(g.edges)
        rng.shuffle(e)
        nn = n
        out = e[len(e) // 4:]
        for (a, b) in e[:len(e) // 4]:
            k = rng.randint(1, 4)
            prev = a
            for _ in range(k):
                out.append((prev, nn))
                prev = nn
                nn += 1
            out.append((prev, b))
        for _ in range(rng.randint(0, 6)):
            out.append((rng.randrange(nn), nn))
            nn += 1
        rng.shuffle(out)
        out = [(a, b) if rng.random() < 0.5 else (b, a) for a, b in out]
        yield f"random{t}", nn, out
    for t in range(2):
        yield f"multibcc{t}", *multi_block(rng)


def multi_block(rng):
    """Several biconnected blocks (a Hamiltonian cycle plus random chords)
    sharing cut vertices along a random tree, a quarter of the edges
    subdivided into chains of 1-3 vertices, pendant vertices, ids shuffled."""
    edges, nn, verts = [], 0, []
    for b in range(rng.randint(4, 6)):
        k = rng.randint(5, 14)
        ids = list(range(nn, nn + k))
        if verts:
            ids[0] = rng.choice(verts)  # share a cut vertex with an earlier block
            ids[1:] = list(range(nn, nn + k - 1))
            nn += k - 1
        else:
            nn += k
        blk = [(ids[i], ids[(i + 1) % k]) for i in range(k)]
        blk += [(ids[i], ids[j]) for i in range(k) for j in range(i + 2, k)
                if (i, j) != (0, k - 1) and rng.random() < 0.25]
        edges += blk
        verts += ids
    rng.shuffle(edges)
    out = []
    for i, (a, b) in enumerate(edges):
        if i % 4 == 0:
            prev = a
            for _ in range(rng.randint(1, 3)):
                out.append((prev, nn))
                prev = nn
                nn += 1
            out.append((prev, b))
        else:
            out.append((a, b))
    for _ in range(rng.randint(2, 6)):
        out.append((rng.randrange(nn), nn))
        nn += 1
    perm = list(range(nn))
    rng.shuffle(perm)
    out = [(perm[a], perm[b]) for a, b in out]
    rng.shuffle(out)
    return nn, out


def reference_structure(n, edges):
    """The reference's structure output for an input, restated here
    independently of the library (pure Python, run only where the goldens
    are made):
      * BCC edge lists in emission order: recursive edge-stack Tarjan over
        adjacency lists in file order (Codes/graph.cpp:52-57), from vertex 0
        and then every unvisited vertex (Codes/bicc.cpp:39-44), a block
        popped when dfs[u] <= low[v] after the return from v
        (Codes/bicc.cpp:81-161);
      * local ids: first appearance over each popped edge list, u before v
        (Codes/modified_apsp.cpp:336-357);
      * stdout: ' n_bcc= N' (Codes/modified_apsp.cpp:271); per block of more
        than two vertices, in emission order, 'BCC number #: i nodes n edges
        2|E| Total BCC's are N' (Codes/to_eargraph_construct.cpp:54-65);
        'active_vert_count A free_vert_count F ' (Codes/main.cpp:50), A the
        distinct endpoints of chains with at least one interior vertex (empty
        ear pieces are never emitted, Codes/bcc_bfs.cpp:109-113;
        Codes/readeardata.cpp:188) and F the vertices that are neither an
        endpoint nor a chain vertex (Codes/to_eargraph_construct.cpp:114-120,
        162), summed over the blocks of more than two vertices."""
    import sys
    sys.setrecursionlimit(max(10000, 4 * n))
    adj = [[] for _ in range(n)]
    for u, v in edges:
        adj[u].append(v)
        adj[v].append(u)
    dfs, low = [-1] * n, [-1] * n
    stack, blocks = [], []
    clock = [0]

    def visit(u, parent):
        clock[0] += 1
        dfs[u] = low[u] = clock[0]
        for v in adj[u]:
            if dfs[v] == -1:
                stack.append((u, v))
                visit(v, u)
                low[u] = min(low[u], low[v])
                if dfs[u] <= low[v] and stack:
                    blk = []
                    while stack:
                        e = stack.pop()
                        blk.append(e)
                        if e == (u, v):
                            break
                    blocks.append(blk)
            elif dfs[v] < dfs[u] and v != parent:
                stack.append((u, v))
                low[u] = min(low[u], dfs[v])

    for r in range(n):
        if dfs[r] == -1:
            visit(r, -1)
    nb = len(blocks)
    lines = [f" n_bcc= {nb}"]
    local_ids, active, free, per_block = [], 0, 0, []
    for i, blk in enumerate(blocks):
        ids = []
        for u, v in blk:
            for x in (u, v):
                if x not in ids:
                    ids.append(x)
        local_ids.append(ids)
        if len(ids) <= 2:
            continue
        lines.append(f"BCC number #: {i} nodes {len(ids)} edges {2 * len(blk)} Total BCC's are {nb}")
        deg = {x: 0 for x in ids}
        nbr = {x: [] for x in ids}
        for u, v in blk:
            deg[u] += 1
            deg[v] += 1
            nbr[u].append(v)
            nbr[v].append(u)
        if max(deg.values()) >= 3:
            chain = {x for x in ids if deg[x] == 2}
            ends = {y for x in chain for y in nbr[x] if deg[y] >= 3}
        else:
            chain, ends = set(), set()  # a cycle: no vertex of degree 3 splits it
        active += len(ends)
        free += len(ids) - len(ends) - len(chain)
        per_block.append([i, len(ids) - len(ends) - len(chain), len(ends)])
    lines.append(f"active_vert_count {active} free_vert_count {free} ")
    return {"lines": lines, "bcc_sizes": [[len(ids), len(b)] for ids, b in zip(local_ids, blocks)],
            "local_ids": local_ids, "blocks": per_block}


def main():
    res = {}
    for name, n, edges in graphs():
        G = nx.Graph()
        G.add_nodes_from(range(n))
        G.add_edges_from(edges)
        bc = nx.betweenness_centrality(G, normalized=False)
        bccs = [sorted(sorted(e) for e in comp) for comp in nx.biconnected_component_edges(G)]
        bccs.sort()
        art = sorted(nx.articulation_points(G))
        s = max(range(n), key=lambda v: (G.degree(v), -v))
        lv = nx.single_source_shortest_path_length(G, s)
        sig = {}
        for t in range(n):
            sig[t] = 1 if t == s else sum(1 for _ in nx.all_shortest_paths(G, s, t))
        res[name] = {
            "n": n, "edges": [list(e) for e in edges],
            "bc": [2.0 * bc[v] for v in range(n)],
            "bccs": bccs, "articulation": art,
            "source": s, "levels": [lv[v] for v in range(n)], "sigma": [sig[v] for v in range(n)],
            "reference_structure": reference_structure(n, edges),
        }
    with open(os.path.join(HERE, "goldens.json"), "w") as f:
        json.dump(res, f, indent=0, sort_keys=True)
    print("wrote", len(res), "graphs")


if __name__ == "__main__":
    main()
