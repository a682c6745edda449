"""bcx_bc_small_noalloc (the batched small-component run with every scratch
buffer in a static device array) against the allocating entry point it
varies, bcx_run on the batched path, through the C-ABI.

Two calls in a row per precision, the second on a smaller component with a
source subset; the first component is a 45x45 grid whose corner path counts
pass 2^64, so its tiles also take the FP64 rerun.  Bars: FP64 within 1e-9
relative of bcx_run's FP64 result, FP32-compensated within 1e-6 relative of
it; zero entries exactly zero.  Device buffers of the inputs and the output
are allocated here, by the test, as a caller would.
"""
import ctypes as C

import numpy as np
import pytest

import bcx
import oracle
from helpers import relerr

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


class DevArray:
    """A device copy of a numpy array (hipMalloc by the caller, not the library)."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(max(self.a.nbytes, 8))) == 0
        assert hip().hipMemcpy(self.p, self.a.ctypes.data_as(C.c_void_p), C.c_size_t(self.a.nbytes), 1) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def noalloc(engine, c, srcs, prec):
    bufs = [DevArray(c.row_ptr.astype(np.int32)), DevArray(c.col.astype(np.int32)),
            DevArray(c.reach.astype(np.int32))]
    s = DevArray(np.asarray(srcs, np.int32)) if srcs is not None else None
    out = DevArray(np.zeros(c.n, np.float64))
    try:
        rc = bcx.lib().bcx_bc_small_noalloc(engine.h, c.n, c.arcs, bufs[0].p, bufs[1].p, bufs[2].p,
                                            s.p if s else None, len(srcs) if srcs is not None else 0,
                                            prec, out.p)
        if rc:
            raise bcx.BcxError(rc, bcx.lib().bcx_last_error(engine.h).decode())
        return out.get()
    finally:
        for b in bufs + [out] + ([s] if s else []):
            b.free()


def grid_edges(w, h):
    e = [(y * w + x, y * w + x + 1) for y in range(h) for x in range(w - 1)]
    e += [(y * w + x, (y + 1) * w + x) for y in range(h - 1) for x in range(w)]
    return e


def components():
    big = bcx.Host.from_edges(45 * 45, grid_edges(45, 45)).decompose().components()[0]
    n, uv = bcx.generate(1, 0, 0.2)
    small = max(bcx.Host.from_edges(n, uv).decompose().components(), key=lambda c: c.n)
    assert small.n < big.n <= 2048
    return big, small


@pytest.mark.parametrize("prec", [bcx.PREC_FP64, bcx.PREC_FP32C])
def test_noalloc_matches_batched_run(engine, prec):
    big, small = components()
    rng = np.random.default_rng(11)
    sub = np.sort(rng.choice(small.n, size=small.n // 3, replace=False)).astype(np.int32)
    for c, srcs in ((big, None), (small, sub)):
        g = engine.upload([c])
        try:
            want, st = g.run(path=bcx.PATH_BATCHED, precision=bcx.PREC_FP64,
                             subset=None if srcs is None else srcs.astype(np.int64))
        finally:
            g.free()
        got = noalloc(engine, c, srcs, prec)
        tol = 1e-9 if prec == bcx.PREC_FP64 else 1e-6
        mx, mean, zeros = relerr(got, want)
        print(f"no-alloc n {c.n} sources {c.n if srcs is None else len(srcs)} prec {prec} "
              f"max {mx:.2e} f64 tiles in the run {st['n_f64_sigma']}", flush=True)
        assert mx < tol and zeros, (c.n, prec, mx)
        ref = oracle.component_bc(c.row_ptr, c.col, c.reach, sources=srcs)
        mx, _, zeros = relerr(got, ref)
        assert mx < tol and zeros
    assert st is not None


def test_noalloc_rejects_what_it_cannot_hold(engine):
    big, small = components()
    cyc = bcx.Host.from_edges(2100, [(i, (i + 1) % 2100) for i in range(2100)]).decompose().components()[0]
    assert cyc.n == 2100
    with pytest.raises(bcx.BcxError) as e:
        noalloc(engine, cyc, None, bcx.PREC_FP64)
    assert e.value.code == -1 and "BCX_NOALLOC_MAX_N" in str(e.value)
    with pytest.raises(bcx.BcxError) as e:
        noalloc(engine, small, np.array([0, small.n], np.int32), bcx.PREC_FP64)
    assert e.value.code == -1 and "out of range" in str(e.value)
    # the context and the scratch array stay usable after a rejected call
    got = noalloc(engine, small, None, bcx.PREC_FP64)
    mx, _, zeros = relerr(got, oracle.component_bc(small.row_ptr, small.col, small.reach))
    assert mx < 1e-9 and zeros
