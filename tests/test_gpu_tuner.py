"""The batched path's schedule tuner and every schedule it can choose.

A component of at least 16384 vertices gets its kernel (workgroup or
multi-launch tiles, up to 131072 vertices) and its multi-launch schedule
(level-kernel blocks per CU, tiles in flight, chain contraction) measured
once on sample tiles; later runs reuse the choice.  Each choice the tuner can
make is also forced through the options and checked against the oracle.
"""
import numpy as np
import pytest

import bcx
import oracle
from helpers import relerr

pytestmark = pytest.mark.gpu
B = bcx.PATH_BATCHED
MULTI = bcx.FLAG_NO_SMALL_KERNEL | bcx.FLAG_NO_WG_KERNEL


def giant(cfg, scale):
    n, uv = bcx.generate(cfg, 0, scale)
    comps = bcx.Host.from_edges(n, uv).decompose().components()
    return max(comps, key=lambda c: c.n)


def grid_edges(w, h):
    e = [(y * w + x, y * w + x + 1) for y in range(h) for x in range(w - 1)]
    e += [(y * w + x, (y + 1) * w + x) for y in range(h - 1) for x in range(w)]
    return e


def sample_bc(engine, c, k, seed, **kw):
    """BC of c over k random sources (batched) and the oracle's."""
    rng = np.random.default_rng(seed)
    srcs = np.sort(rng.choice(c.n, size=k, replace=False))
    g = engine.upload([c])
    try:
        got, st = g.run(subset=srcs, path=B, **kw)
    finally:
        g.free()
    want = oracle.component_bc(c.row_ptr, c.col, c.reach, sources=srcs.astype(np.int32), threads=16)
    return got, want, st


def test_tuned_once_and_cached(engine):
    """The first run of a 33k-vertex mixed component on the multi-launch
    tiles measures their schedule, the second reuses it; both equal the
    oracle."""
    c = giant(2, 0.04)
    assert c.n >= 16384
    g = engine.upload([c])
    try:
        rng = np.random.default_rng(3)
        srcs = np.sort(rng.choice(c.n, size=512, replace=False))
        kw = dict(subset=srcs, path=B, precision=bcx.PREC_FP64, flags=bcx.FLAG_NO_WG_KERNEL)
        got1, st1 = g.run(**kw)
        got2, st2 = g.run(**kw)
    finally:
        g.free()
    print("case n", c.n, "tuned", st1["n_tuned"], st2["n_tuned"], "bpc", st1["sched_blocks_per_cu"],
          "slots", st1["sched_slots"], "contracted", st1["n_contracted"], st2["n_contracted"], flush=True)
    assert st1["n_tuned"] == 1 and st2["n_tuned"] == 0
    assert st1["n_comp_multi"] == 1
    assert st1["sched_blocks_per_cu"] in (4, 8, 16) and st1["sched_slots"] in (4, 8)
    assert (st1["sched_blocks_per_cu"], st1["sched_slots"]) == (st2["sched_blocks_per_cu"], st2["sched_slots"])
    assert st1["n_contracted"] == st2["n_contracted"]
    want = oracle.component_bc(c.row_ptr, c.col, c.reach, sources=srcs.astype(np.int32), threads=16)
    for got in (got1, got2):
        mx, _, zeros = relerr(got, want)
        assert mx < 1e-9 and zeros


def test_kernel_measured_for_mid_size_component(engine):
    """With the kernel free, the same component is measured on the workgroup
    kernel and the multi-launch tiles; whichever wins, the result is exact."""
    c = giant(2, 0.04)
    got, want, st = sample_bc(engine, c, 256, 4, precision=bcx.PREC_FP64)
    print("case n", c.n, "tuned", st["n_tuned"], "wg", st["n_comp_wg"], "multi", st["n_comp_multi"], flush=True)
    assert st["n_tuned"] == 1 and st["n_comp_wg"] + st["n_comp_multi"] == 1
    mx, _, zeros = relerr(got, want)
    assert mx < 1e-9 and zeros


@pytest.mark.parametrize("kw", [dict(blocks_per_cu=2), dict(blocks_per_cu=16), dict(batch_width=1),
                                dict(batch_width=3), dict(flags=bcx.FLAG_NO_CONTRACT),
                                dict(flags=bcx.FLAG_NO_TUNE)])
def test_forced_schedules(engine, kw):
    """Every multi-launch schedule the tuner chooses from, forced; the run
    reports the forced values and equals the oracle."""
    c = giant(5, 0.04)
    flags = MULTI | kw.pop("flags", 0)
    got, want, st = sample_bc(engine, c, 256, 5, precision=bcx.PREC_FP32C, flags=flags, **kw)
    print("case", kw, flags, "n", c.n, "bpc", st["sched_blocks_per_cu"], "slots", st["sched_slots"], flush=True)
    if "blocks_per_cu" in kw:
        assert st["sched_blocks_per_cu"] == kw["blocks_per_cu"]
    if "batch_width" in kw:
        assert st["sched_slots"] == kw["batch_width"]
    if flags & bcx.FLAG_NO_CONTRACT:
        assert st["n_contracted"] == 0
    if flags & bcx.FLAG_NO_TUNE:
        assert st["n_tuned"] == 0
    mx, _, zeros = relerr(got, want)
    assert mx < 1e-6 and zeros


def test_kernel_measured_for_deep_component(engine):
    """A 160x160 grid (25.6k vertices, 319 BFS levels) is measured on both
    kernels; the workgroup kernel wins by a wide margin, and the result is
    exact either way."""
    (c,) = bcx.Host.from_edges(160 * 160, grid_edges(160, 160)).decompose().components()
    got, want, st = sample_bc(engine, c, 192, 11, precision=bcx.PREC_FP64)
    print("case grid160 tuned", st["n_tuned"], "wg", st["n_comp_wg"], "multi", st["n_comp_multi"], flush=True)
    assert st["n_tuned"] == 1 and st["n_comp_wg"] == 1
    mx, _, zeros = relerr(got, want)
    assert mx < 1e-9 and zeros
