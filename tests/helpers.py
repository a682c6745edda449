"""Shared test helpers (pure Python / numpy)."""
import itertools
from collections import deque

import numpy as np


def relerr(a, b):
    """max and mean relative error over b > 0, and whether b == 0 entries are exact."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    pos = b > 0
    r = np.abs(a[pos] - b[pos]) / b[pos] if pos.any() else np.zeros(1)
    return float(r.max()), float(r.mean()), bool(np.all(a[~pos] == 0))


def brute_bc(n, edges):
    """Ordered-pair BC by enumerating every shortest path (tiny graphs only)."""
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    bc = np.zeros(n)
    for s in range(n):
        dist = [-1] * n
        dist[s] = 0
        q = deque([s])
        while q:
            v = q.popleft()
            for w in adj[v]:
                if dist[w] < 0:
                    dist[w] = dist[v] + 1
                    q.append(w)
        for t in range(n):
            if t == s:
                continue
            paths = []

            def walk(p):
                v = p[-1]
                if v == t:
                    paths.append(list(p))
                    return
                for w in adj[v]:
                    if dist[w] == dist[v] + 1 and dist[w] <= dist[t]:
                        p.append(w)
                        walk(p)
                        p.pop()

            walk([s])
            for p in paths:
                for v in p[1:-1]:
                    bc[v] += 1.0 / len(paths)
    return bc


def comb(a, b):
    from math import comb as c
    return c(a, b)
