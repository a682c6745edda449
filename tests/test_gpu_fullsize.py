"""Parity of the batched path on the benchmark workloads at their stated size.

Workloads (generator seed 0, scale 1.0):
  2  1M-vertex mixed graph (Chung-Lu core + degree-2 chains + pendant trees)
  3  519k-vertex Chung-Lu block + 40,000 small blocks (12.8M vertices)
  4  256x256 grid - 5% of edges + 2,000 chains (FP64 path counts, ~1e150)
  5  R-MAT scale 20 + chains
Every run uses default flags, so the kernel, tile order, hub threshold,
relabelled copy and slots in flight are the production choices.

Oracle coverage (the oracle cannot run every source of workloads 2, 3, 5 in
a test): workload 4 all sources, whole graph; workloads 2 and 5 a stratified
sample of 16 whole tuned tiles (1,024 sources: every (T/16)-th tile of the
library's tile order); workload 3 eight whole tiles of its large block plus
every source of 200 random small blocks, in one run.  Bars: BC max relative
error 1e-6 (FP32-compensated) and 1e-9 (FP64), zero-BC vertices exactly zero;
one whole 64-lane tile per workload is also compared lane by lane (levels and
path counts bit-exact, dependencies as in test_gpu_tiles).
"""
import numpy as np
import pytest

import bcx
import oracle
from helpers import relerr
from tilecheck import check_tile

pytestmark = pytest.mark.gpu

_cache = {}


def workload(cfg):
    if cfg not in _cache:
        _cache.clear()  # one full-size workload in memory at a time
        n, uv = bcx.generate(cfg, 0, 1.0)
        h = bcx.Host.from_edges(n, uv).decompose()
        _cache[cfg] = (n, uv, h, h.components())
    return _cache[cfg]


def whole_tiles(order, k):
    """k evenly spaced whole tiles of a tile order."""
    T = len(order) // 64
    picks = sorted({(i * T) // k for i in range(k)})
    return [order[t * 64:(t + 1) * 64] for t in picks]


def check_run(g, ci, c, subset_local, want, tag, plain=True):
    sub = np.sort(np.asarray(subset_local, np.int64)) + g.offsets[ci]
    runs = [(bcx.PATH_BATCHED, bcx.PREC_FP32C, 1e-6), (bcx.PATH_BATCHED, bcx.PREC_FP64, 1e-9)]
    if plain:
        runs.append((bcx.PATH_PLAIN, bcx.PREC_FP64, 1e-9))
    for path, prec, tol in runs:
        got, st = g.run(path=path, precision=prec, subset=sub)
        mx, mean, zeros = relerr(got[g.offsets[ci]:g.offsets[ci] + c.n], want)
        print(f"{tag} path {path} prec {prec} sources {len(sub)} max {mx:.2e} mean {mean:.2e} "
              f"f64 {st['n_f64_sigma']} t {st['t_kernel_s']:.3f}s", flush=True)
        assert mx < tol and zeros, (tag, path, prec, mx)


@pytest.mark.parametrize("cfg", [2, 5])
def test_power_law_core_tiles(engine, cfg):
    n, uv, h, comps = workload(cfg)
    ci = int(np.argmax([c.n for c in comps]))
    c = comps[ci]
    print(f"workload {cfg}: n {n} m {len(uv)} core {c.n} arcs {c.arcs}", flush=True)
    g = engine.upload(comps)
    try:
        order = g.batched_order(ci)
        tiles = whole_tiles(order, 16)
        sample = np.concatenate(tiles).astype(np.int32)
        want = oracle.component_bc(c.row_ptr, c.col, c.reach, sources=np.sort(sample), threads=16)
        check_run(g, ci, c, sample, want, f"workload {cfg}")
    finally:
        g.free()
    # lane-by-lane: the first tile (around the highest-degree vertex) and a middle one
    for t in (0, len(tiles) // 2):
        check_tile(engine, c, tiles[t], bcx.PREC_FP32C, tag=f"workload {cfg} tile {t}",
                   want_kernel=bcx.KERNEL_MULTI)
    # plain path: a few sources, levels and counts bit-exact
    rng = np.random.default_rng(cfg)
    for s in rng.choice(c.n, size=3, replace=False):
        lv, su, sf, dl, isf = engine.debug_source(c, int(s), use_chains=True)
        olv, osu, osf, odl, ovf = oracle.source(c.row_ptr, c.col, c.reach, int(s))
        np.testing.assert_array_equal(lv, olv)
        np.testing.assert_array_equal(su if not isf else sf, osu if not isf else osf)
        np.testing.assert_allclose(dl, odl, rtol=1e-12, atol=0)


def test_grid_all_sources(engine):
    """Workload 4, every source, whole graph (cut term included)."""
    n, uv, h, comps = workload(4)
    rp, col = oracle.csr_from_edges(n, uv)
    want = oracle.brandes_bc(rp, col, threads=16)
    for prec, tol in ((bcx.PREC_FP32C, 1e-6), (bcx.PREC_FP64, 1e-9)):
        got, st = engine.host_bc(h, path=bcx.PATH_BATCHED, precision=prec)
        mx, mean, zeros = relerr(got, want)
        print(f"workload 4 prec {prec} max {mx:.2e} mean {mean:.2e} f64 {st['n_f64_sigma']} "
              f"depth {st['max_level']} t {st['t_kernel_s']:.3f}s", flush=True)
        assert st["n_f64_sigma"] > 0
        assert mx < tol and zeros, (prec, mx)
    ci = int(np.argmax([c.n for c in comps]))
    c = comps[ci]
    g = engine.upload(comps)
    try:
        order = g.batched_order(ci)
    finally:
        g.free()
    for t in (0, len(order) // 128):
        check_tile(engine, c, order[t * 64:(t + 1) * 64], bcx.PREC_FP32C, tag=f"workload 4 tile {t}",
                   want_kernel=bcx.KERNEL_WG, want_f64=True)


def test_many_components(engine):
    """Workload 3: eight whole tiles of the large block and every source of
    200 random small blocks in one run (the small-component kernel, the
    workgroup kernel where chosen and the multi-launch tiles together)."""
    n, uv, h, comps = workload(3)
    sizes = np.array([c.n for c in comps])
    ci = int(np.argmax(sizes))
    c = comps[ci]
    rng = np.random.default_rng(3)
    small = rng.choice(np.nonzero(sizes <= 2048)[0], size=200, replace=False)
    print(f"workload 3: n {n} components {len(comps)} core {c.n} small sizes {sizes[small].min()}-"
          f"{sizes[small].max()}", flush=True)
    g = engine.upload(comps)
    try:
        order = g.batched_order(ci)
        tiles = whole_tiles(order, 8)
        core_src = np.sort(np.concatenate(tiles)).astype(np.int32)
        sub = [core_src.astype(np.int64) + g.offsets[ci]]
        sub += [np.arange(comps[k].n, dtype=np.int64) + g.offsets[k] for k in small]
        sub = np.sort(np.concatenate(sub))
        want_core = oracle.component_bc(c.row_ptr, c.col, c.reach, sources=core_src, threads=16)
        want_small = {int(k): oracle.component_bc(comps[k].row_ptr, comps[k].col, comps[k].reach)
                      for k in small}
        for prec, tol in ((bcx.PREC_FP32C, 1e-6), (bcx.PREC_FP64, 1e-9)):
            got, st = g.run(path=bcx.PATH_BATCHED, precision=prec, subset=sub)
            mx, mean, zeros = relerr(got[g.offsets[ci]:g.offsets[ci] + c.n], want_core)
            print(f"workload 3 core prec {prec} max {mx:.2e} mean {mean:.2e}", flush=True)
            assert mx < tol and zeros, ("core", prec, mx)
            worst = 0.0
            for k, w in want_small.items():
                mx, _, zeros = relerr(got[g.offsets[k]:g.offsets[k] + comps[k].n], w)
                assert zeros, ("small", k, prec)
                worst = max(worst, mx)
            print(f"workload 3 small blocks prec {prec} max {worst:.2e}", flush=True)
            assert worst < tol, ("small", prec, worst)
    finally:
        g.free()
    check_tile(engine, c, tiles[0], bcx.PREC_FP32C, tag="workload 3 core tile 0", want_kernel=bcx.KERNEL_MULTI)
    k = int(small[np.argmax(sizes[small])])
    check_tile(engine, comps[k], np.arange(min(64, comps[k].n)), bcx.PREC_FP32C,
               tag=f"workload 3 block {k}", want_kernel=bcx.KERNEL_SMALL)
