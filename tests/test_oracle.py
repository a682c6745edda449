"""Pins the CPU oracle before anything is compared against it.

* orc_brandes_bc vs 2 x networkx goldens (tests/golden/goldens.json), brute
  force path enumeration and closed forms (path, star, cycle, complete graph)
  plus the checksum sum_v BC(v) = sum_{s != t} (d(s,t) - 1);
* orc_source levels/sigma vs networkx goldens and closed forms (grid
  binomials, complete graph, cycle), and the uint64 overflow flag;
* orc_chain_source vs a direct orc_source from the same vertex;
* a planted mistake in the oracle's inputs (a dropped reach weight) is caught.
"""
import numpy as np
import pytest

import oracle
from helpers import brute_bc, comb, relerr


def csr(n, edges):
    return oracle.csr_from_edges(n, np.asarray(edges, np.int64).reshape(-1, 2))


def test_brandes_vs_networkx_goldens(goldens):
    for name, g in goldens.items():
        rp, col = csr(g["n"], g["edges"])
        bc = oracle.brandes_bc(rp, col, threads=2)
        np.testing.assert_allclose(bc, g["bc"], rtol=1e-12, atol=1e-9, err_msg=name)


def test_brandes_vs_bruteforce():
    rng = np.random.default_rng(7)
    for t in range(6):
        n = int(rng.integers(5, 11))
        edges = [(i, i + 1) for i in range(n - 1)]  # connected
        for a in range(n):
            for b in range(a + 2, n):
                if rng.random() < 0.3:
                    edges.append((a, b))
        rp, col = csr(n, edges)
        np.testing.assert_allclose(oracle.brandes_bc(rp, col), brute_bc(n, edges), rtol=1e-12, atol=1e-12)


def test_closed_forms():
    n = 9
    rp, col = csr(n, [(i, i + 1) for i in range(n - 1)])
    np.testing.assert_allclose(oracle.brandes_bc(rp, col), [2 * i * (n - 1 - i) for i in range(n)])
    rp, col = csr(n, [(0, i) for i in range(1, n)])
    want = np.zeros(n)
    want[0] = (n - 1) * (n - 2)
    np.testing.assert_allclose(oracle.brandes_bc(rp, col), want)
    rp, col = csr(7, [(i, j) for i in range(7) for j in range(i + 1, 7)])
    np.testing.assert_array_equal(oracle.brandes_bc(rp, col), np.zeros(7))


def test_checksum_distance_sum():
    rng = np.random.default_rng(3)
    n = 60
    edges = [(i, int(rng.integers(0, i))) for i in range(1, n)]
    edges += [(int(a), int(b)) for a, b in rng.integers(0, n, (80, 2)) if a != b]
    edges = sorted({(min(a, b), max(a, b)) for a, b in edges})
    rp, col = csr(n, edges)
    total = 0
    for s in range(n):
        lv, *_ = oracle.source(rp, col, None, s)
        total += int(np.sum(np.maximum(lv - 1, 0)))
    assert abs(oracle.brandes_bc(rp, col).sum() - total) < 1e-7 * total


def test_source_levels_sigma_goldens(goldens):
    for name, g in goldens.items():
        rp, col = csr(g["n"], g["edges"])
        lv, su, sf, dl, ovf = oracle.source(rp, col, None, g["source"])
        assert not ovf
        np.testing.assert_array_equal(lv, g["levels"], err_msg=name)
        np.testing.assert_array_equal(su, np.asarray(g["sigma"], np.uint64), err_msg=name)
        np.testing.assert_array_equal(sf, np.asarray(g["sigma"], np.float64), err_msg=name)


def grid_edges(w, h):
    e = []
    for y in range(h):
        for x in range(w):
            v = y * w + x
            if x + 1 < w:
                e.append((v, v + 1))
            if y + 1 < h:
                e.append((v, v + w))
    return e


def test_sigma_grid_binomial_and_overflow():
    # corner source of a w x h grid: sigma(x, y) = C(x + y, x)
    w = h = 40
    rp, col = csr(w * h, grid_edges(w, h))
    # entries below 2^64 stay exact even though the far corner overflows
    lv, su, sf, dl, ovf = oracle.source(rp, col, None, 0)
    for y in range(0, h, 7):
        for x in range(0, w, 5):
            want = comb(x + y, x)
            assert lv[y * w + x] == x + y
            if want < 2 ** 53:
                assert sf[y * w + x] == float(want)
            if want < 2 ** 64:
                assert int(su[y * w + x]) == want
    assert ovf and comb(78, 39) > 2 ** 64


def test_sigma_overflow_flag():
    w = h = 40
    rp, col = csr(w * h, grid_edges(w, h))
    *_, ovf = oracle.source(rp, col, None, 0)
    assert ovf
    w = h = 20  # C(38,19) ~ 3.5e10 fits
    rp, col = csr(w * h, grid_edges(w, h))
    *_, ovf = oracle.source(rp, col, None, 0)
    assert not ovf


def test_sources_rows_equal_single_sources():
    """oracle.sources (threaded, many sources) gives row by row exactly what
    oracle.source gives, including the overflow flag and the FP64 shadow;
    a grid corner (overflow) and an interior vertex (no overflow) included."""
    w = h = 40
    rp, col = csr(w * h, grid_edges(w, h))
    reach = np.arange(w * h, dtype=np.int32) % 3
    srcs = np.array([0, 820, 5, 1599, 41], np.int32)
    lv, su, sf, dl, ovf = oracle.sources(rp, col, reach, srcs, threads=3)
    for i, s in enumerate(srcs):
        a = oracle.source(rp, col, reach, int(s))
        np.testing.assert_array_equal(lv[i], a[0])
        np.testing.assert_array_equal(su[i], a[1])
        np.testing.assert_array_equal(sf[i].view(np.uint64), a[2].view(np.uint64))
        np.testing.assert_array_equal(dl[i].view(np.uint64), a[3].view(np.uint64))
        assert ovf[i] == a[4]
    assert ovf[0] and not ovf[1]


def test_sigma_complete_and_cycle():
    rp, col = csr(6, [(i, j) for i in range(6) for j in range(i + 1, 6)])
    lv, su, *_ = oracle.source(rp, col, None, 2)
    assert list(su) == [1] * 6
    n = 10
    rp, col = csr(n, [(i, (i + 1) % n) for i in range(n)])
    lv, su, *_ = oracle.source(rp, col, None, 0)
    assert su[5] == 2 and su[4] == 1 and lv[5] == 5


def test_chain_source_equals_direct_source(goldens):
    import bcx
    for name in ("sample_test1", "sample_test4", "random0", "random1", "random2", "random3", "barbell5_3"):
        g = goldens[name]
        h = bcx.Host.from_edges(g["n"], g["edges"]).decompose()
        for comp in h.components():
            for k in range(comp.n_chains):
                nodes = comp.chain_nodes[comp.chain_ptr[k]:comp.chain_ptr[k + 1]]
                L, R = comp.chain_ends[2 * k], comp.chain_ends[2 * k + 1]
                for pos in range(len(nodes)):
                    a = oracle.chain_source(comp.row_ptr, comp.col, comp.reach, L, R, nodes, pos)
                    b = oracle.source(comp.row_ptr, comp.col, comp.reach, int(nodes[pos]))
                    np.testing.assert_array_equal(a[0], b[0], err_msg=name)
                    np.testing.assert_array_equal(a[1], b[1], err_msg=name)
                    np.testing.assert_allclose(a[3], b[3], rtol=1e-13, err_msg=name)


def test_planted_mistake_is_caught(goldens):
    """Zeroing the reach weights (a dropped term) must break whole-graph parity."""
    import bcx
    g = goldens["sample_test4"]
    h = bcx.Host.from_edges(g["n"], g["edges"]).decompose()
    bc = h.cut_term().astype(float)
    for c in h.components():
        np.add.at(bc, c.gid, oracle.component_bc(c.row_ptr, c.col, np.zeros_like(c.reach)))
    assert relerr(bc, g["bc"])[0] > 1e-3
