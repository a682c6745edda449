"""GPU parity of the batched multi-source path through the C-ABI.

Bar: levels and sigma bit-exact against the oracle (uint64, and FP64 sigma in
the oracle's CSR pull order); with FP64 dependency terms delta within 1e-12
and BC within 1e-9 relative; with FP32-compensated terms delta within 2e-6
and BC within 1e-6 max relative (the FP32C bar); zero BC exact.
"""
import numpy as np
import pytest

import bcx
import oracle
from helpers import relerr

pytestmark = pytest.mark.gpu
B = bcx.PATH_BATCHED


def comps_of(n, uv, chains=True):
    return bcx.Host.from_edges(n, uv).decompose(use_chains=chains).components()


def check_source(engine, c, s, prec, **kw):
    lv, su, sf, dl, isf = engine.debug_source(c, s, path=B, precision=prec, **kw)
    olv, osu, osf, odl, ovf = oracle.source(c.row_ptr, c.col, c.reach, s)
    np.testing.assert_array_equal(lv, olv)
    if isf:
        np.testing.assert_array_equal(sf, osf)
    else:
        assert not ovf
        np.testing.assert_array_equal(su, osu)
    tol = 2e-6 if prec == bcx.PREC_FP32C else 1e-12
    np.testing.assert_allclose(dl, odl, rtol=tol, atol=1e-300)


@pytest.mark.parametrize("prec", [bcx.PREC_FP64, bcx.PREC_FP32C])
def test_levels_sigma_delta_goldens(engine, goldens, prec):
    for name, g in goldens.items():
        print("case", name, prec, flush=True)
        for c in comps_of(g["n"], g["edges"], chains=False):
            for s in range(0, c.n, max(1, c.n // 7)):
                check_source(engine, c, s, prec)


@pytest.mark.parametrize("cfg,scale", [(1, 0.25), (2, 0.003), (4, 0.05), (5, 0.003)])
def test_sampled_sources_generated(engine, cfg, scale):
    n, uv = bcx.generate(cfg, 0, scale)
    big = max(comps_of(n, uv), key=lambda c: c.n)
    rng = np.random.default_rng(cfg)
    print("case cfg", cfg, "n", big.n, flush=True)
    for s in rng.choice(big.n, size=min(8, big.n), replace=False):
        check_source(engine, big, int(s), bcx.PREC_FP32C)


@pytest.mark.parametrize("prec,tol", [(bcx.PREC_FP64, 1e-9), (bcx.PREC_FP32C, 1e-6)])
def test_component_bc_goldens(engine, goldens, prec, tol):
    for name, g in goldens.items():
        comps = comps_of(g["n"], g["edges"])
        if not comps:
            continue
        got, st = engine.bc(comps, path=B, precision=prec)
        for c, b in zip(comps, got):
            mx, _, zeros = relerr(b, oracle.component_bc(c.row_ptr, c.col, c.reach))
            assert mx < tol and zeros, name


@pytest.mark.parametrize("cfg,scale", [(1, 1.0), (2, 0.003), (4, 0.1), (5, 0.003)])
@pytest.mark.parametrize("prec,tol", [(bcx.PREC_FP64, 1e-9), (bcx.PREC_FP32C, 1e-6)])
def test_whole_graph_bc_generated(engine, cfg, scale, prec, tol):
    n, uv = bcx.generate(cfg, 0, scale)
    h = bcx.Host.from_edges(n, uv).decompose()
    rp, col = oracle.csr_from_edges(n, uv)
    want = oracle.brandes_bc(rp, col, threads=16)
    print("case cfg", cfg, "n", n, "prec", prec, flush=True)
    bc, st = engine.host_bc(h, path=B, precision=prec)
    mx, mean, zeros = relerr(bc, want)
    print("maxrel", mx, "mean", mean, flush=True)
    assert mx < tol and zeros


def test_fp64_sigma_grid(engine):
    w = 40
    edges = []
    for y in range(w):
        for x in range(w):
            v = y * w + x
            if x + 1 < w:
                edges.append((v, v + 1))
            if y + 1 < w:
                edges.append((v, v + w))
    c = comps_of(w * w, edges)[0]
    for s in (0, w - 1, w * w // 2 + 3):
        check_source(engine, c, s, bcx.PREC_FP64)
    got, st = engine.bc([c], path=B, precision=bcx.PREC_FP64)
    assert st["n_f64_sigma"] > 0
    mx, _, zeros = relerr(got[0], oracle.component_bc(c.row_ptr, c.col, c.reach))
    assert mx < 1e-9 and zeros


@pytest.mark.parametrize("prec", [bcx.PREC_FP64, bcx.PREC_FP32C])
def test_sigma_word_escalation_through_hub_slices(engine, prec):
    """A hub of degree k = 520 (above the smallest hub threshold, 512 arcs,
    so its sigma is summed by arc slices with atomics) reaches sigma = k^4 =
    7.3e10 > 2^32 from vertex 0 through four complete bipartite layers: the
    tile leaves 32-bit words at the hub's level, its forward stops there and
    it is rerun with 64-bit words.
    Layers: 0 - L1 = L2 = L3 = L4 - hub (k vertices per layer)."""
    k = 520
    layers = [list(range(1 + i * k, 1 + (i + 1) * k)) for i in range(4)]
    hub = 1 + 4 * k
    edges = [(0, a) for a in layers[0]]
    for i in range(3):
        edges += [(a, b) for a in layers[i] for b in layers[i + 1]]
    edges += [(d, hub) for d in layers[3]]
    c = comps_of(hub + 1, edges)[0]
    assert c.n == hub + 1
    gid = list(c.gid)
    _, su, _, _, _ = oracle.source(c.row_ptr, c.col, c.reach, gid.index(0))
    assert int(su[gid.index(hub)]) == k ** 4 > 2 ** 32
    flags = bcx.FLAG_NO_SMALL_KERNEL | bcx.FLAG_NO_WG_KERNEL
    for s in (0, hub, 5, 3 * k):
        check_source(engine, c, gid.index(s), prec, flags=flags)
    want = oracle.component_bc(c.row_ptr, c.col, c.reach)
    got, _ = engine.bc([c], path=B, precision=prec, flags=flags)
    mx, _, zeros = relerr(got[0], want)
    assert mx < (1e-6 if prec == bcx.PREC_FP32C else 1e-9) and zeros


def test_tile_order_tuner(engine):
    """The tile-order tuner keeps ball tiles for a grid and BFS-rank tiles for
    a power-law core (whichever needs fewer (vertex, level) entries on sample
    tiles); either way the BC is the oracle's."""
    w = 60
    edges = [(y * w + x, y * w + x + 1) for y in range(w) for x in range(w - 1)]
    edges += [(y * w + x, (y + 1) * w + x) for y in range(w - 1) for x in range(w)]
    grid = comps_of(w * w, edges)[0]
    n, uv = bcx.generate(2, 0, 0.01)
    core = max(comps_of(n, uv), key=lambda c: c.n)
    assert core.n > 4000
    for c, balls, flags in ((grid, 1, 0), (core, 0, bcx.FLAG_WG_KERNEL), (core, 0, bcx.FLAG_NO_WG_KERNEL)):
        got, st = engine.bc([c], path=B, precision=bcx.PREC_FP64, flags=flags)
        assert st["n_ball_tiles"] == balls, (c.n, flags)
        mx, _, zeros = relerr(got[0], oracle.component_bc(c.row_ptr, c.col, c.reach))
        assert mx < 1e-9 and zeros


def test_batched_order_whole_tiles(engine):
    """bcx_batched_order: every component's vertices in the tuned tile order
    (a permutation); a subset made of whole tiles of that order gives the
    oracle's partial BC for those sources."""
    n, uv = bcx.generate(2, 0, 0.03)
    comps = comps_of(n, uv)
    g = engine.upload(comps)
    try:
        for ci, c in enumerate(comps):
            o = g.batched_order(ci)
            assert np.array_equal(np.sort(o), np.arange(c.n)), ci
        ci = int(np.argmax([c.n for c in comps]))
        c = comps[ci]
        o = g.batched_order(ci)
        T = c.n // 64
        tiles = np.concatenate([o[t * 64:(t + 1) * 64] for t in (0, T // 2, T - 1)])
        sub = np.sort(tiles).astype(np.int64) + g.offsets[ci]
        want = oracle.component_bc(c.row_ptr, c.col, c.reach, sources=np.sort(tiles).astype(np.int32))
        for prec, tol in ((bcx.PREC_FP64, 1e-9), (bcx.PREC_FP32C, 1e-6)):
            got, _ = g.run(path=B, precision=prec, subset=sub)
            mx, _, zeros = relerr(got[g.offsets[ci]:g.offsets[ci] + c.n], want)
            assert mx < tol and zeros, prec
    finally:
        g.free()


def test_threshold_hook_and_subset(engine):
    n, uv = bcx.generate(4, 0, 0.05)
    c = max(comps_of(n, uv), key=lambda c: c.n)
    got, st = engine.bc([c], path=B, precision=bcx.PREC_FP32C, sigma_threshold_log2=6)
    assert st["n_f64_sigma"] > 0
    mx, _, zeros = relerr(got[0], oracle.component_bc(c.row_ptr, c.col, c.reach))
    assert mx < 1e-6 and zeros
    g = engine.upload([c])
    try:
        sub = np.sort(np.random.default_rng(5).choice(c.n, size=c.n // 3, replace=False))
        got, st = g.run(path=B, precision=bcx.PREC_FP32C, subset=sub)
        mx, _, zeros = relerr(got, oracle.component_bc(c.row_ptr, c.col, c.reach, sources=sub))
        assert mx < 1e-6 and zeros
    finally:
        g.free()


@pytest.mark.parametrize("flags", [0, bcx.FLAG_NO_SMALL_KERNEL])
def test_many_small_components(engine, flags):
    """Workload 3 shape (a large BCC plus many small ones) at small scale:
    the one-launch small-component kernel and the multi-launch tiles agree
    with the oracle."""
    n, uv = bcx.generate(3, 0, 0.002)
    h = bcx.Host.from_edges(n, uv).decompose()
    comps = h.components()
    assert sum(c.n <= 2048 for c in comps) > 20
    rp, col = oracle.csr_from_edges(n, uv)
    want = oracle.brandes_bc(rp, col, threads=16)
    for prec, tol in ((bcx.PREC_FP64, 1e-9), (bcx.PREC_FP32C, 1e-6)):
        bc, st = engine.host_bc(h, path=B, precision=prec, flags=flags)
        mx, _, zeros = relerr(bc, want)
        print("flags", flags, "prec", prec, "maxrel", mx, st["n_launches"], flush=True)
        assert mx < tol and zeros


def test_small_kernel_fp64_sigma_rerun(engine):
    """A small grid whose sigma passes a forced 2^6 threshold: the tiles are
    listed and rerun with FP64 sigma inside the small-component kernel."""
    w = 20
    edges = [(y * w + x, y * w + x + 1) for y in range(w) for x in range(w - 1)]
    edges += [(y * w + x, (y + 1) * w + x) for y in range(w - 1) for x in range(w)]
    c = comps_of(w * w, edges)[0]
    got, st = engine.bc([c], path=B, precision=bcx.PREC_FP32C, sigma_threshold_log2=6)
    assert st["n_f64_sigma"] > 0
    mx, _, zeros = relerr(got[0], oracle.component_bc(c.row_ptr, c.col, c.reach))
    assert mx < 1e-6 and zeros


def test_small_kernel_graph_replays(engine):
    """The captured small-component launch replayed several times, with eager
    work and synchronisations in between, matches the eager launch every time."""
    n, uv = bcx.generate(3, 0, 0.001)
    comps = [c for c in bcx.Host.from_edges(n, uv).decompose().components() if c.n <= 2048]
    g = engine.upload(comps)
    try:
        eager, _ = g.run(path=B, precision=bcx.PREC_FP64, flags=bcx.FLAG_NO_GRAPH)
        for rep in range(4):
            got, _ = g.run(path=B, precision=bcx.PREC_FP64)
            np.testing.assert_allclose(got, eager, rtol=1e-12, atol=0, err_msg=f"replay {rep}")
            other, _ = g.run(path=bcx.PATH_PLAIN, precision=bcx.PREC_FP64)  # eager work between replays
            np.testing.assert_allclose(other, eager, rtol=1e-9, atol=0)
    finally:
        g.free()


@pytest.mark.parametrize("kernel_flags", [bcx.FLAG_NO_SMALL_KERNEL | bcx.FLAG_NO_WG_KERNEL,
                                          bcx.FLAG_NO_SMALL_KERNEL | bcx.FLAG_WG_KERNEL])
def test_tile_graph_replays(engine, kernel_flags):
    """The captured level launches of multi-launch tiles (one graph per slot,
    count mode and level count) and the captured workgroup-kernel pass,
    replayed over several runs with eager work and synchronisations between
    them, equal the eager launches every time and the oracle."""
    n, uv = bcx.generate(2, 0, 0.01)
    core = max(comps_of(n, uv), key=lambda c: c.n)
    want = oracle.component_bc(core.row_ptr, core.col, core.reach)
    g = engine.upload([core])
    try:
        eager, st = g.run(path=B, precision=bcx.PREC_FP64, flags=kernel_flags | bcx.FLAG_NO_GRAPH)
        assert st["n_graph_launches"] == 0
        mx, _, zeros = relerr(eager, want)
        assert mx < 1e-9 and zeros
        for rep in range(4):
            got, st = g.run(path=B, precision=bcx.PREC_FP64, flags=kernel_flags)
            assert st["n_graph_launches"] > 0, rep
            np.testing.assert_allclose(got, eager, rtol=1e-12, atol=0, err_msg=f"replay {rep}")
            other, _ = g.run(path=bcx.PATH_PLAIN, precision=bcx.PREC_FP64)  # eager work between replays
            np.testing.assert_allclose(other, eager, rtol=1e-9, atol=0)
        got, st = g.run(path=B, precision=bcx.PREC_FP32C, flags=kernel_flags)
        mx, _, zeros = relerr(got, want)
        assert mx < 1e-6 and zeros
    finally:
        g.free()


@pytest.mark.parametrize("flags", [0, bcx.FLAG_NO_SMALL_KERNEL])
def test_fp32c_sigma_beyond_f32_range(engine, flags):
    """Layers of 3 vertices, consecutive layers completely joined: sigma
    reaches 3^99 (> 1e47), beyond f32 range, so 1/sigma underflows f32.  The
    FP64-sigma tiles of the FP32-compensated mode keep BC within 1e-6, both in
    the small-component kernel and in the multi-launch tiles."""
    L, k = 100, 3
    edges = [(i * k + a, (i + 1) * k + b) for i in range(L - 1) for a in range(k) for b in range(k)]
    c = comps_of(L * k, edges, chains=False)[0]
    assert c.n == L * k
    got, st = engine.bc([c], path=B, precision=bcx.PREC_FP32C, flags=flags)
    assert st["n_f64_sigma"] > 0
    want = oracle.component_bc(c.row_ptr, c.col, c.reach)
    mx, _, zeros = relerr(got[0], want)
    print("flags", flags, "maxrel", mx, flush=True)
    assert mx < 1e-6 and zeros
    for s in (0, k * L // 2):
        check_source(engine, c, s, bcx.PREC_FP32C)


def grid_edges(w, h, base=0):
    e = [(base + y * w + x, base + y * w + x + 1) for y in range(h) for x in range(w - 1)]
    e += [(base + y * w + x, base + (y + 1) * w + x) for y in range(h - 1) for x in range(w)]
    return e


@pytest.mark.parametrize("spl", [0, 2, 4])
@pytest.mark.parametrize("prec,tol", [(bcx.PREC_FP64, 1e-9), (bcx.PREC_FP32C, 1e-6)])
def test_wg_kernel_deep_components(engine, prec, tol, spl):
    """Two grids (60x60 and 45x50, sigma past 2^64 in both) joined by a
    bridge: both biconnected components run in one workgroup-per-tile launch
    (uint64 pass, then the FP64 rerun of the overflowed tiles) and match the
    oracle and the multi-launch tiles; with tiles of 64, 128 and 256
    sources."""
    e = grid_edges(60, 60) + grid_edges(45, 50, base=3600) + [(3599, 3600)]
    n = 3600 + 45 * 50
    h = bcx.Host.from_edges(n, e).decompose()
    comps = h.components()
    assert sorted(c.n for c in comps) == [2250, 3600]
    rp, col = oracle.csr_from_edges(n, e)
    want = oracle.brandes_bc(rp, col, threads=16)
    got, st = engine.host_bc(h, path=B, precision=prec, flags=bcx.FLAG_WG_KERNEL, sources_per_lane=spl)
    assert st["n_f64_sigma"] > 0
    mx, mean, zeros = relerr(got, want)
    print("wg prec", prec, "spl", spl, "maxrel", mx, "launches", st["n_launches"], flush=True)
    assert mx < tol and zeros
    ref, _ = engine.host_bc(h, path=B, precision=prec, flags=bcx.FLAG_NO_WG_KERNEL)
    mx2, _, _ = relerr(got, ref)
    assert mx2 < 2 * tol


@pytest.mark.parametrize("spl", [0, 2, 4])
@pytest.mark.parametrize("cfg,scale", [(2, 0.003), (4, 0.1), (5, 0.004)])
def test_wg_kernel_generated(engine, cfg, scale, spl):
    """Forced onto the workgroup kernel, generated graphs (a power-law core
    with hubs, the chained grid, R-MAT) match the oracle, with tiles of 64,
    128 and 256 sources."""
    n, uv = bcx.generate(cfg, 0, scale)
    h = bcx.Host.from_edges(n, uv).decompose()
    assert max(c.n for c in h.components()) > 2048
    rp, col = oracle.csr_from_edges(n, uv)
    want = oracle.brandes_bc(rp, col, threads=16)
    for prec, tol in ((bcx.PREC_FP64, 1e-9), (bcx.PREC_FP32C, 1e-6)):
        got, st = engine.host_bc(h, path=B, precision=prec, flags=bcx.FLAG_WG_KERNEL, sources_per_lane=spl)
        mx, mean, zeros = relerr(got, want)
        print("cfg", cfg, "spl", spl, "prec", prec, "maxrel", mx, "launches", st["n_launches"], flush=True)
        assert mx < tol and zeros


def test_wg_kernel_threshold_and_subset(engine):
    """Forced 2^6 sigma threshold (FP64 rerun of every tile) and a source
    subset on the workgroup kernel."""
    c = comps_of(3600, grid_edges(60, 60))[0]
    got, st = engine.bc([c], path=B, precision=bcx.PREC_FP32C, sigma_threshold_log2=6,
                        flags=bcx.FLAG_WG_KERNEL)
    assert st["n_f64_sigma"] == c.n
    mx, _, zeros = relerr(got[0], oracle.component_bc(c.row_ptr, c.col, c.reach))
    assert mx < 1e-6 and zeros
    g = engine.upload([c])
    try:
        sub = np.sort(np.random.default_rng(7).choice(c.n, size=500, replace=False))
        got, st = g.run(path=B, precision=bcx.PREC_FP64, subset=sub, flags=bcx.FLAG_WG_KERNEL)
        mx, _, zeros = relerr(got, oracle.component_bc(c.row_ptr, c.col, c.reach, sources=sub))
        assert mx < 1e-9 and zeros
    finally:
        g.free()


def test_failed_run_leaves_context_clean():
    """A multi-launch run that fails part-way -- the device-memory cap is
    reached while a tile's level lists grow (ball tiles of a 60x60 grid hold
    ~14 list entries per vertex, the lists start at 8) -- abandons the tiles
    in flight on the other slots.  The next run on the same context, on a
    smaller component, must still equal the oracle: no level-mask bits of the
    abandoned forwards and no unfolded partial BC may survive."""
    flags = bcx.FLAG_NO_SMALL_KERNEL | bcx.FLAG_NO_WG_KERNEL
    grid = comps_of(3600, grid_edges(60, 60))[0]
    n, uv = bcx.generate(2, 0, 0.003)
    core = max(comps_of(n, uv), key=lambda c: c.n)
    assert 2048 < core.n < grid.n
    want = oracle.component_bc(core.row_ptr, core.col, core.reach)

    def attempt(cap):
        try:
            eng = bcx.Engine(0, mem_cap=cap)
        except bcx.BcxError:
            return None, None
        try:
            eng.bc([grid], path=B, precision=bcx.PREC_FP64, flags=flags)
            return eng, ""
        except bcx.BcxError as e:
            return eng, str(e)

    lo, hi = 1 << 20, 1 << 30  # the grid fails at lo and runs at hi
    while hi - lo > 4096:
        mid = (lo + hi) // 2
        eng, msg = attempt(mid)
        if eng is not None:
            eng.close()
        if msg == "":
            hi = mid
        else:
            lo = mid
    print("smallest cap that runs the grid:", hi, flush=True)
    for k in range(1, 33):
        eng, msg = attempt(hi - 4096 * k)
        if eng is None:
            continue
        if "growth" not in msg:
            eng.close()
            continue
        try:
            print("failed as wanted:", msg, flush=True)
            for prec, tol in ((bcx.PREC_FP64, 1e-9), (bcx.PREC_FP32C, 1e-6)):
                got, st = eng.bc([core], path=B, precision=prec, flags=flags)
                mx, _, zeros = relerr(got[0], want)
                assert mx < tol and zeros, (prec, mx)
        finally:
            eng.close()
        return
    pytest.fail("no memory cap made the list growth fail")
