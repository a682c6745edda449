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
        }
    with open(os.path.join(HERE, "goldens.json"), "w") as f:
        json.dump(res, f, indent=0, sort_keys=True)
    print("wrote", len(res), "graphs")


if __name__ == "__main__":
    main()
