"""Per-lane parity of whole batched tiles on every kernel of the batched path.

bcx_debug_tile runs up to 64 sources as the lanes of one tile on the kernel a
run would pick (or the one the flags force) and returns every lane's levels,
path counts and dependencies.  Each lane is compared with the oracle's
single-source result (oracle.sources):
  * levels: bit-exact (int32);
  * path counts: bit-exact -- integers when the tile's counts fit 64 bits,
    else FP64 summed in the oracle's CSR pull order;
  * dependencies: FP64 terms within 1e-12 relative; FP32-compensated terms
    within 1e-5 relative (one f32 rounding of the term per level; the BC bar
    of that mode stays 1e-6, checked in the run tests);
  * exact zeros where the oracle's dependency is zero.
"""
import numpy as np
import pytest

import bcx
import oracle
from tilecheck import check_tile

pytestmark = pytest.mark.gpu

SMALL = 0
WG = bcx.FLAG_NO_SMALL_KERNEL | bcx.FLAG_WG_KERNEL
MULTI = bcx.FLAG_NO_SMALL_KERNEL | bcx.FLAG_NO_WG_KERNEL
KERNELS = {"small": (SMALL, bcx.KERNEL_SMALL), "wg": (WG, bcx.KERNEL_WG), "multi": (MULTI, bcx.KERNEL_MULTI),
           "multi64": (MULTI, bcx.KERNEL_MULTI), "multi128": (MULTI, bcx.KERNEL_MULTI),
           "wg128": (WG, bcx.KERNEL_WG), "wg256": (WG, bcx.KERNEL_WG)}
# sources per lane of the multi-launch and workgroup tiles ("multi", "wg": the defaults)
SPL = {"multi64": 1, "multi128": 2, "wg128": 2, "wg256": 4}


def comps_of(n, uv, chains=True):
    return bcx.Host.from_edges(n, uv).decompose(use_chains=chains).components()


def grid_edges(w, h, base=0):
    e = [(base + y * w + x, base + y * w + x + 1) for y in range(h) for x in range(w - 1)]
    e += [(base + y * w + x, base + (y + 1) * w + x) for y in range(h - 1) for x in range(w)]
    return e


@pytest.mark.parametrize("prec", [bcx.PREC_FP64, bcx.PREC_FP32C])
@pytest.mark.parametrize("kernel", ["small", "wg", "multi"])
def test_golden_tiles(engine, goldens, prec, kernel):
    """The first 64 vertices (or all) of every golden component as one tile."""
    flags, kid = KERNELS[kernel]
    for name, g in goldens.items():
        for c in comps_of(g["n"], g["edges"], chains=False):
            srcs = np.arange(min(64, c.n))[::-1]  # lanes in reverse id order
            check_tile(engine, c, srcs, prec, flags, kid, tag=f"{name}/{kernel}", spl=SPL.get(kernel, 0))


@pytest.mark.parametrize("spl", [2, 4])
@pytest.mark.parametrize("cfg,scale", [(1, 0.1), (2, 0.002), (4, 0.02), (5, 0.002)])
@pytest.mark.parametrize("kernel", ["multi", "wg"])
def test_generated_wide_tiles(engine, cfg, scale, spl, kernel):
    """Whole tiles of 128 or 256 sources on the multi-launch and workgroup
    kernels (two or four sources per lane): every source against the oracle."""
    flags, kid = KERNELS[kernel]
    tw = 64 * spl
    n, uv = bcx.generate(cfg, 0, scale)
    comps = comps_of(n, uv)
    c = comps[int(np.argmax([c.n for c in comps]))]
    g = engine.upload([c])
    try:
        order = g.batched_order(0)
    finally:
        g.free()
    T = (c.n + tw - 1) // tw
    for t in sorted({0, T // 2, T - 1}):
        srcs = order[t * tw:(t + 1) * tw]
        check_tile(engine, c, srcs, bcx.PREC_FP32C, flags, kid, tag=f"cfg{cfg}/w{t}", spl=spl)
    check_tile(engine, c, order[:tw], bcx.PREC_FP64, flags, kid, tag=f"cfg{cfg}/w0/fp64", spl=spl)


@pytest.mark.parametrize("kernel", ["small", "wg", "multi", "multi64", "multi128", "wg128", "wg256"])
@pytest.mark.parametrize("cfg,scale", [(1, 0.1), (2, 0.002), (4, 0.02), (5, 0.002)])
def test_generated_tiles(engine, cfg, scale, kernel):
    """Whole tuned tiles (bcx_batched_order) of the largest component of each
    generated workload, scaled so every kernel can take it."""
    flags, kid = KERNELS[kernel]
    n, uv = bcx.generate(cfg, 0, scale)
    comps = comps_of(n, uv)
    ci = int(np.argmax([c.n for c in comps]))
    c = comps[ci]
    if kernel == "small" and c.n > 2048:
        pytest.skip("component above the small kernel's size")
    g = engine.upload(comps)
    try:
        order = g.batched_order(ci)
    finally:
        g.free()
    T = (c.n + 63) // 64
    for t in sorted({0, T // 2, T - 1}):
        srcs = order[t * 64:(t + 1) * 64]
        check_tile(engine, c, srcs, bcx.PREC_FP32C, flags, kid, tag=f"cfg{cfg}/t{t}/{kernel}", spl=SPL.get(kernel, 0))


@pytest.mark.parametrize("kernel", ["small", "wg", "multi", "multi64", "multi128", "wg128", "wg256"])
def test_fp64_counts_grid_tile(engine, kernel):
    """40x40 grid: corner path counts C(78,39) > 2^64, so a tile holding a
    corner runs with FP64 counts on every kernel, bit-exact in pull order."""
    flags, kid = KERNELS[kernel]
    c = comps_of(1600, grid_edges(40, 40))[0]
    gid = list(c.gid)
    srcs = [gid.index(v) for v in (0, 39, 1560, 1599)] + [gid.index(v) for v in range(41, 41 + 60)]
    spl = SPL.get(kernel, 0)
    check_tile(engine, c, srcs, bcx.PREC_FP64, flags, kid, want_f64=True, tag=f"grid40/{kernel}", spl=spl)
    check_tile(engine, c, srcs, bcx.PREC_FP32C, flags, kid, want_f64=True, tag=f"grid40/{kernel}/fp32c", spl=spl)
    if kernel in ("multi128", "wg128"):
        wide = srcs + [v for v in range(400, 464)]
        check_tile(engine, c, wide, bcx.PREC_FP32C, flags, kid, want_f64=True, tag=f"grid40/{kernel}/128", spl=2)
    if kernel == "wg256":
        wide = srcs + [v for v in range(400, 592)]
        check_tile(engine, c, wide, bcx.PREC_FP64, flags, kid, want_f64=True, tag="grid40/wg/256", spl=4)
    if kernel == "multi":
        wide = srcs + [v for v in range(400, 592)]
        check_tile(engine, c, wide, bcx.PREC_FP32C, flags, kid, want_f64=True, tag="grid40/multi/256")


@pytest.mark.parametrize("kernel", ["wg", "multi"])
def test_hub_slices_tile(engine, kernel):
    """A hub of degree 520 (sliced by the multi-launch kernels, whose smallest
    hub threshold is 512 arcs) whose path count from vertex 0 is 520^4 > 2^32:
    the tile's 32-bit counts are escalated to 64 bits; all 64 lanes exact."""
    flags, kid = KERNELS[kernel]
    k = 520
    layers = [list(range(1 + i * k, 1 + (i + 1) * k)) for i in range(4)]
    hub = 1 + 4 * k
    edges = [(0, a) for a in layers[0]]
    for i in range(3):
        edges += [(a, b) for a in layers[i] for b in layers[i + 1]]
    edges += [(d, hub) for d in layers[3]]
    c = comps_of(hub + 1, edges)[0]
    gid = list(c.gid)
    srcs = [gid.index(0), gid.index(hub)] + [gid.index(v) for v in range(5, 5 * 63, 5)][:62]
    for prec in (bcx.PREC_FP64, bcx.PREC_FP32C):
        check_tile(engine, c, srcs, prec, flags, kid, want_f64=False, tag=f"hub/{kernel}/{prec}")


@pytest.mark.parametrize("kernel", ["small", "wg", "multi"])
def test_counts_beyond_f32_tile(engine, kernel):
    """Layers of 3 vertices, consecutive layers complete: counts up to 3^99;
    the FP32-compensated mode runs such tiles with FP64 counts and terms."""
    flags, kid = KERNELS[kernel]
    L, k = 100, 3
    edges = [(i * k + a, (i + 1) * k + b) for i in range(L - 1) for a in range(k) for b in range(k)]
    c = comps_of(L * k, edges, chains=False)[0]
    srcs = np.arange(0, c.n, 5)[:64]
    check_tile(engine, c, srcs, bcx.PREC_FP32C, flags, kid, want_f64=True, tag=f"layers/{kernel}")


def test_tile_threshold_hook(engine):
    """A forced 2^6 count threshold moves a tile to FP64 on every kernel;
    every lane still equals the oracle bit for bit."""
    c = comps_of(400, grid_edges(20, 20))[0]
    srcs = np.arange(64)
    for kernel, (flags, kid) in KERNELS.items():
        lv, su, sf, dl, isf, kern = engine.debug_tile(c, srcs, precision=bcx.PREC_FP32C, flags=flags,
                                                      sigma_threshold_log2=6)
        assert isf and kern == kid, kernel
        olv, _, osf, odl, _ = oracle.sources(c.row_ptr, c.col, c.reach, srcs)
        np.testing.assert_array_equal(lv, olv)
        np.testing.assert_array_equal(sf.view(np.uint64), osf.view(np.uint64))
        np.testing.assert_allclose(dl, odl, rtol=1e-12, atol=0)


def test_tile_probe_rejects_bad_sources(engine):
    c = comps_of(400, grid_edges(20, 20))[0]
    for srcs in ([0, 0], [-1], [400], list(range(65))):
        with pytest.raises(bcx.BcxError):
            engine.debug_tile(c, srcs)
