"""Lane-by-lane comparison of a batched tile (bcx_debug_tile) with the
oracle; shared by the GPU tile and full-size tests."""
import numpy as np

import bcx
import oracle


def check_tile(engine, c, srcs, prec, flags=0, want_kernel=None, want_f64=None, tag="", spl=0):
    """Runs the tile (spl: sources per lane of the multi-launch tiles, 0 =
    default) and compares every source with the oracle; returns the max
    relative dependency error."""
    srcs = np.asarray(srcs, np.int32)
    lv, su, sf, dl, isf, kern = engine.debug_tile(c, srcs, precision=prec, flags=flags, sources_per_lane=spl)
    if want_kernel is not None:
        assert kern == want_kernel, (tag, kern)
    if want_f64 is not None:
        assert isf == want_f64, (tag, isf)
    olv, osu, osf, odl, ovf = oracle.sources(c.row_ptr, c.col, c.reach, srcs, threads=16)
    bad = np.nonzero((lv != olv).any(axis=1))[0]
    assert len(bad) == 0, f"{tag}: levels differ on lanes {bad[:8]}"
    if isf:
        bad = np.nonzero((sf.view(np.uint64) != osf.view(np.uint64)).any(axis=1))[0]
        assert len(bad) == 0, f"{tag}: FP64 path counts differ on lanes {bad[:8]}"
    else:
        assert not ovf.any(), tag
        bad = np.nonzero((su != osu).any(axis=1))[0]
        assert len(bad) == 0, f"{tag}: path counts differ on lanes {bad[:8]}"
    zero = odl == 0
    assert np.all(dl[zero] == 0), f"{tag}: non-zero dependency where the oracle has zero"
    err = float(np.max(np.abs(dl[~zero] - odl[~zero]) / odl[~zero])) if (~zero).any() else 0.0
    tol = 1e-5 if (prec == bcx.PREC_FP32C and not isf) else 1e-12
    print(f"tile {tag} kernel {kern} lanes {len(srcs)} f64 {isf} max dep err {err:.2e}", flush=True)
    assert err <= tol, (tag, err)
    return err
