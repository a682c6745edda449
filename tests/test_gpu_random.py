"""Seeded random graphs of many shapes through every GPU kernel choice.

Each graph (connected, simple, unit weight) is decomposed on the host and its
whole-graph BC is computed by the plain path (FP64) and by the batched path
with each kernel forced in turn: the small-component kernel, the
workgroup-per-tile kernel and the multi-launch tiles, in FP64 and in
FP32-compensated precision, and compared with the oracle's FP64 Brandes on
the whole graph.  Bars: FP64 1e-9, FP32C 1e-6 max relative error, zero BC exact.
"""
import numpy as np
import pytest

import bcx
import oracle
from helpers import relerr

pytestmark = pytest.mark.gpu

B = bcx.PATH_BATCHED
KERNELS = {
    "small": 0,
    "workgroup": bcx.FLAG_NO_SMALL_KERNEL | bcx.FLAG_WG_KERNEL,
    "multi-launch": bcx.FLAG_NO_SMALL_KERNEL | bcx.FLAG_NO_WG_KERNEL,
}


def connect(n, edges, rng):
    """Adds edges until the graph is connected (union-find over the given edges)."""
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in edges:
        parent[find(a)] = find(b)
    roots = sorted({find(v) for v in range(n)})
    for r0, r1 in zip(roots, roots[1:]):
        a = int(rng.choice([v for v in range(n) if find(v) == find(r0)]))
        b = int(rng.choice([v for v in range(n) if find(v) == find(r1)]))
        edges.append((a, b))
        parent[find(a)] = find(b)
    return edges


def shape(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "tree+chords":
        n = int(rng.integers(40, 300))
        e = [(v, int(rng.integers(0, v))) for v in range(1, n)]
        e += [tuple(sorted(map(int, rng.choice(n, 2, replace=False)))) for _ in range(n // 4)]
    elif kind == "cycles":
        n, e, base = 0, [], 0
        for _ in range(int(rng.integers(3, 8))):
            k = int(rng.integers(3, 30))
            e += [(base + i, base + (i + 1) % k) for i in range(k)]
            if base:
                e.append((base - 1, base))  # bridge between consecutive cycles
            base += k
        n = base
    elif kind == "grid-holes":
        w, h = int(rng.integers(8, 20)), int(rng.integers(8, 20))
        n = w * h
        e = [(y * w + x, y * w + x + 1) for y in range(h) for x in range(w - 1)]
        e += [(y * w + x, (y + 1) * w + x) for y in range(h - 1) for x in range(w)]
        keep = rng.random(len(e)) > 0.15
        e = [p for p, k in zip(e, keep) if k]
    elif kind == "power-law":
        n = int(rng.integers(100, 400))
        w = 1.0 / np.arange(1, n + 1) ** 0.8
        p = w / w.sum()
        e = []
        for _ in range(3 * n):
            a, b = map(int, rng.choice(n, 2, p=p))
            if a != b:
                e.append((min(a, b), max(a, b)))
    elif kind == "chained-cliques":
        n, e, base = 0, [], 0
        for _ in range(int(rng.integers(3, 7))):
            k = int(rng.integers(4, 9))
            e += [(base + i, base + j) for i in range(k) for j in range(i + 1, k)]
            if base:
                L = int(rng.integers(1, 6))  # a chain of L vertices to the previous clique
                prev = base - 1
                for c in range(L):
                    e.append((prev, base + k + c))
                    prev = base + k + c
                e.append((prev, base))
                base += L
            base += k
        n = base
    else:
        raise ValueError(kind)
    e = sorted({(min(a, b), max(a, b)) for a, b in e if a != b})
    e = connect(n, e, rng)
    return n, sorted({(min(a, b), max(a, b)) for a, b in e})


CASES = [(k, s) for k in ("tree+chords", "cycles", "grid-holes", "power-law", "chained-cliques") for s in (1, 2)]


@pytest.mark.parametrize("kind,seed", CASES)
def test_random_graph_all_kernels(engine, kind, seed):
    n, edges = shape(kind, seed)
    h = bcx.Host.from_edges(n, edges).decompose()
    rp, col = oracle.csr_from_edges(n, edges)
    want = oracle.brandes_bc(rp, col, threads=16)
    print("case", kind, seed, "n", n, "m", len(edges), flush=True)
    got, _ = engine.host_bc(h, path=bcx.PATH_PLAIN, precision=bcx.PREC_FP64)
    mx, _, zeros = relerr(got, want)
    assert mx < 1e-9 and zeros, ("plain", mx)
    for name, flags in KERNELS.items():
        for prec, tol in ((bcx.PREC_FP64, 1e-9), (bcx.PREC_FP32C, 1e-6)):
            got, _ = engine.host_bc(h, path=B, precision=prec, flags=flags)
            mx, _, zeros = relerr(got, want)
            assert mx < tol and zeros, (name, prec, mx)
