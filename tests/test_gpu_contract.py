"""Contracted chains on the multi-launch tiles.

A chain of K arcs between L and R that is longer than a path joining L and R
outside every chain is never used by a shortest path between two vertices off
the chain; the multi-launch tiles take such chains out of the tile graph and
give their interiors' levels, path counts and dependencies in closed form from
the endpoints', and run a source inside one as a lane seeded at both ends.
Every lane of such tiles is compared with the oracle to the tile bar (levels
and counts bit-exact, dependencies 1e-12 / 1e-5), whole runs with the oracle
and with BCX_FLAG_NO_CONTRACT (every chain vertex in the tile graph).
"""
import numpy as np
import pytest

import bcx
import oracle
from helpers import relerr
from tilecheck import check_tile

pytestmark = pytest.mark.gpu
B = bcx.PATH_BATCHED
MULTI = bcx.FLAG_NO_SMALL_KERNEL | bcx.FLAG_NO_WG_KERNEL


def comps_of(n, uv):
    return bcx.Host.from_edges(n, uv).decompose(use_chains=True).components()


def theta_graph():
    """K6 with six chains between its vertices: two parallel ones (0-1), one
    of a single interior (2-3), odd and even lengths, three chains at vertex
    0.  Every chain is longer than the one arc joining its ends."""
    e = [(a, b) for a in range(6) for b in range(a + 1, 6)]
    nxt = 6
    for L, R, k in ((0, 1, 5), (0, 1, 3), (2, 3, 1), (0, 2, 2), (4, 5, 8), (0, 3, 4)):
        path = [L] + list(range(nxt, nxt + k)) + [R]
        nxt += k
        e += list(zip(path[:-1], path[1:]))
    return nxt, e


def contracted(engine, c):
    """Chain interiors a multi-launch run of component c takes out."""
    _, st = engine.bc([c], path=B, flags=MULTI)
    return st["n_contracted"]


def chain_rich_tiles(c, rng, width=64):
    """Tiles of `width` sources: chain interiors only, and interiors mixed
    with their endpoints and other vertices."""
    nodes = np.asarray(c.chain_nodes)
    ends = np.unique(np.asarray(c.chain_ends))
    t1 = rng.choice(nodes, size=min(width, len(nodes)), replace=False)
    mix = np.concatenate([rng.choice(nodes, size=min(3 * width // 8, len(nodes)), replace=False),
                          rng.choice(ends, size=min(5 * width // 16, len(ends)), replace=False)])
    rest = np.setdiff1d(np.arange(c.n), mix)
    t2 = np.concatenate([mix, rng.choice(rest, size=width - len(mix), replace=False)])
    return [t1, t2]


# tiles of 64, 128 and 256 sources (one, two and four per lane)
WIDTHS = {64: 1, 128: 2, 256: 4}


@pytest.mark.parametrize("width", [64, 128, 256])
@pytest.mark.parametrize("prec", [bcx.PREC_FP64, bcx.PREC_FP32C])
def test_theta_every_source(engine, prec, width):
    """All 29 vertices of the theta graph as one tile: every position of every
    chain is a source once, against the oracle source by source; with wider
    tiles, a tile of another graph that fills the lanes' later words too."""
    n, e = theta_graph()
    (c,) = comps_of(n, e)
    assert c.n_chains == 6
    assert contracted(engine, c) == 23
    spl = WIDTHS[width]
    check_tile(engine, c, np.arange(c.n), prec, MULTI, bcx.KERNEL_MULTI, tag="theta", spl=spl)
    check_tile(engine, c, np.arange(c.n)[::-1], prec, MULTI, bcx.KERNEL_MULTI, tag="theta/rev", spl=spl)
    if width > 64:
        # 64 + 29: lanes 0..28 carry a source in both words
        rng = np.random.default_rng(1)
        big = max(comps_of(*bcx.generate(2, 0, 0.003)), key=lambda c: c.n)
        srcs = np.concatenate([rng.choice(big.n, size=64, replace=False)])
        extra = rng.choice(np.setdiff1d(np.arange(big.n), srcs), width - 35 - 64, replace=False)
        check_tile(engine, big, np.concatenate([srcs, extra]), prec, MULTI, bcx.KERNEL_MULTI,
                   tag=f"mixed{width - 35}", spl=spl)


@pytest.mark.parametrize("width", [64, 128, 256])
@pytest.mark.parametrize("cfg,scale", [(2, 0.003), (5, 0.003), (1, 1.0)])
@pytest.mark.parametrize("prec", [bcx.PREC_FP64, bcx.PREC_FP32C])
def test_contracted_generated_tiles(engine, cfg, scale, prec, width):
    n, uv = bcx.generate(cfg, 0, scale)
    c = max(comps_of(n, uv), key=lambda c: c.n)
    print("case cfg", cfg, "n", c.n, "chains", c.n_chains, "width", width, flush=True)
    assert contracted(engine, c) > 0
    rng = np.random.default_rng(cfg)
    for k, srcs in enumerate(chain_rich_tiles(c, rng, width)):
        check_tile(engine, c, srcs, prec, MULTI, bcx.KERNEL_MULTI, tag=f"cfg{cfg}/t{k}/w{width}", spl=WIDTHS[width])


def test_contracted_tile_fp64_counts(engine):
    """A forced 2^3 count threshold moves the tile to FP64 counts: the chain
    sources' own paths are added at their endpoints in CSR order, so FP64
    counts stay bit-exact."""
    n, uv = bcx.generate(2, 0, 0.003)
    c = max(comps_of(n, uv), key=lambda c: c.n)
    srcs = chain_rich_tiles(c, np.random.default_rng(7), 128)[1]
    lv, su, sf, dl, isf, kern = engine.debug_tile(c, srcs, precision=bcx.PREC_FP64, flags=MULTI,
                                                  sigma_threshold_log2=3)
    assert isf and kern == bcx.KERNEL_MULTI
    olv, _, osf, odl, _ = oracle.sources(c.row_ptr, c.col, c.reach, srcs, threads=16)
    np.testing.assert_array_equal(lv, olv)
    np.testing.assert_array_equal(sf.view(np.uint64), osf.view(np.uint64))
    np.testing.assert_allclose(dl, odl, rtol=1e-12, atol=0)


@pytest.mark.parametrize("cfg,scale", [(2, 0.004), (5, 0.004)])
@pytest.mark.parametrize("prec,tol", [(bcx.PREC_FP64, 1e-9), (bcx.PREC_FP32C, 1e-6)])
def test_contracted_whole_graph(engine, cfg, scale, prec, tol):
    """Whole-graph BC with contraction against the oracle and against the run
    that keeps every chain vertex; the run reports the contracted vertices."""
    n, uv = bcx.generate(cfg, 0, scale)
    h = bcx.Host.from_edges(n, uv).decompose()
    rp, col = oracle.csr_from_edges(n, uv)
    want = oracle.brandes_bc(rp, col, threads=16)
    print("case cfg", cfg, "n", n, "prec", prec, flush=True)
    bc, st = engine.host_bc(h, path=B, precision=prec, flags=MULTI)
    assert st["n_contracted"] > 0
    mx, mean, zeros = relerr(bc, want)
    print("contracted", st["n_contracted"], "maxrel", mx, "mean", mean, flush=True)
    assert mx < tol and zeros
    bc2, st2 = engine.host_bc(h, path=B, precision=prec, flags=MULTI | bcx.FLAG_NO_CONTRACT)
    assert st2["n_contracted"] == 0
    mx2, _, _ = relerr(bc, bc2)
    assert mx2 < 2 * tol


def test_no_contraction_without_chains(engine):
    """Run with use_chains off, nothing is contracted and the result is
    unchanged."""
    n, uv = bcx.generate(2, 0, 0.003)
    h = bcx.Host.from_edges(n, uv).decompose()
    bc1, st1 = engine.host_bc(h, path=B, precision=bcx.PREC_FP64, flags=MULTI)
    bc2, st2 = engine.host_bc(h, path=B, precision=bcx.PREC_FP64, flags=MULTI, use_chains=False)
    assert st1["n_contracted"] > 0 and st2["n_contracted"] == 0
    assert relerr(bc1, bc2)[0] < 1e-9
