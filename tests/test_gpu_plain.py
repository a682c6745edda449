"""GPU parity of the plain per-source path through the C-ABI.

Bar: levels and sigma bit-exact against the oracle (uint64 sigma, and FP64
sigma summed in the oracle's CSR pull order); delta within 1e-12 relative;
BC within 1e-9 relative (FP64 atomics reorder sums), zero-BC vertices exact.
"""
import os
import subprocess

import numpy as np
import pytest

import bcx
import oracle
from helpers import relerr

pytestmark = pytest.mark.gpu

BC_TOL = 1e-9


def comps_of(n, uv, chains=True):
    return bcx.Host.from_edges(n, uv).decompose(use_chains=chains).components()


def check_source(engine, c, s, **kw):
    lv, su, sf, dl, isf = engine.debug_source(c, s, **kw)
    olv, osu, osf, odl, ovf = oracle.source(c.row_ptr, c.col, c.reach, s)
    np.testing.assert_array_equal(lv, olv)
    if isf:
        assert ovf or kw.get("sigma_threshold_log2", 0)
        np.testing.assert_array_equal(sf, osf)  # bit-exact FP64 pull order
    else:
        assert not ovf
        np.testing.assert_array_equal(su, osu)
    np.testing.assert_allclose(dl, odl, rtol=1e-12, atol=1e-300)


def test_levels_sigma_delta_goldens(engine, goldens):
    for name, g in goldens.items():
        print("case", name, flush=True)
        for c in comps_of(g["n"], g["edges"], chains=False):
            for s in range(c.n):
                check_source(engine, c, s, use_chains=False)


def test_chain_sources_match_direct(engine, goldens):
    for name in ("sample_test1", "sample_test2", "sample_test4", "random0", "random1", "random2"):
        g = goldens[name]
        print("case", name, flush=True)
        for c in comps_of(g["n"], g["edges"]):
            for k in range(c.n_chains):
                for x in c.chain_nodes[c.chain_ptr[k]:c.chain_ptr[k + 1]]:
                    check_source(engine, c, int(x), use_chains=True)


@pytest.mark.parametrize("cfg,scale", [(1, 0.25), (2, 0.003), (4, 0.05), (5, 0.003)])
def test_sampled_sources_generated(engine, cfg, scale):
    n, uv = bcx.generate(cfg, 0, scale)
    comps = comps_of(n, uv)
    big = max(comps, key=lambda c: c.n)
    rng = np.random.default_rng(cfg)
    print("case cfg", cfg, "n", big.n, flush=True)
    for s in rng.choice(big.n, size=min(24, big.n), replace=False):
        check_source(engine, big, int(s), use_chains=True)


@pytest.mark.parametrize("chains", [False, True])
def test_component_bc_goldens(engine, goldens, chains):
    for name, g in goldens.items():
        comps = comps_of(g["n"], g["edges"], chains)
        if not comps:
            continue
        got, st = engine.bc(comps, use_chains=chains)
        for c, b in zip(comps, got):
            want = oracle.component_bc(c.row_ptr, c.col, c.reach)
            mx, mean, zeros = relerr(b, want)
            assert mx < BC_TOL and zeros, name


def test_whole_graph_bc_goldens(engine, goldens):
    for name, g in goldens.items():
        h = bcx.Host.from_edges(g["n"], g["edges"]).decompose()
        bc, st = engine.host_bc(h, use_chains=True)
        mx, mean, zeros = relerr(bc, g["bc"])
        assert mx < BC_TOL and zeros, name


@pytest.mark.parametrize("cfg,scale", [(1, 1.0), (2, 0.003), (4, 0.1), (5, 0.003)])
def test_whole_graph_bc_generated(engine, cfg, scale):
    n, uv = bcx.generate(cfg, 0, scale)
    h = bcx.Host.from_edges(n, uv).decompose()
    rp, col = oracle.csr_from_edges(n, uv)
    want = oracle.brandes_bc(rp, col, threads=16)
    print("case cfg", cfg, "n", n, flush=True)
    bc, st = engine.host_bc(h, use_chains=True)
    mx, mean, zeros = relerr(bc, want)
    print("maxrel", mx, "mean", mean, st, flush=True)
    assert mx < BC_TOL and zeros


def test_fp64_sigma_grid(engine):
    """40x40 grid: corner sigma C(78,39) > 2^64 -> FP64 sigma, bit-exact pull order."""
    w = 40
    edges = []
    for y in range(w):
        for x in range(w):
            v = y * w + x
            if x + 1 < w:
                edges.append((v, v + 1))
            if y + 1 < w:
                edges.append((v, v + w))
    comps = comps_of(w * w, edges)
    c = comps[0]
    for s in (0, w - 1, w * w // 2 + 3):
        check_source(engine, c, s, use_chains=False)
    got, st = engine.bc(comps, use_chains=True)
    assert st["n_f64_sigma"] > 0
    mx, _, zeros = relerr(got[0], oracle.component_bc(c.row_ptr, c.col, c.reach))
    assert mx < BC_TOL and zeros


def test_sigma_threshold_hook_forces_fp64(engine):
    n, uv = bcx.generate(4, 0, 0.05)
    comps = comps_of(n, uv)
    c = max(comps, key=lambda c: c.n)
    check_source(engine, c, 0, use_chains=False, sigma_threshold_log2=6)
    got, st = engine.bc([c], use_chains=True, sigma_threshold_log2=6)
    assert st["n_f64_sigma"] > 0
    mx, _, zeros = relerr(got[0], oracle.component_bc(c.row_ptr, c.col, c.reach))
    assert mx < BC_TOL and zeros


def test_subset_and_tiny_row_cap(engine):
    """Source subsets and a 2-row endpoint slab (many chain batches, rows recomputed)."""
    n, uv = bcx.generate(1, 0, 0.5)
    comps = comps_of(n, uv)
    c = max(comps, key=lambda c: c.n)
    g = engine.upload([c])
    try:
        rng = np.random.default_rng(11)
        sub = np.sort(rng.choice(c.n, size=c.n // 3, replace=False))
        got, st = g.run(use_chains=True, subset=sub)
        want = oracle.component_bc(c.row_ptr, c.col, c.reach, sources=sub)
        mx, _, zeros = relerr(got, want)
        assert mx < BC_TOL and zeros
        got, st = g.run(use_chains=True, row_cap_bytes=2 * 12 * (c.n + 64))
        assert st["n_row_only"] > 0
        want = oracle.component_bc(c.row_ptr, c.col, c.reach)
        mx, _, zeros = relerr(got, want)
        assert mx < BC_TOL and zeros
    finally:
        g.free()


def test_memory_cap_is_an_error():
    eng = bcx.Engine(0, mem_cap=1 << 20)
    try:
        n, uv = bcx.generate(1, 0, 0.5)
        comps = comps_of(n, uv)
        with pytest.raises(bcx.BcxError) as e:
            eng.bc(comps)
        assert e.value.code == -4
    finally:
        eng.close()


def test_cli_sample_inputs(goldens, tmp_path):
    exe = os.path.join(bcx.HERE, "betc")
    for k in range(1, 5):
        src = os.path.join(bcx.HERE, "..", "tests", "golden", f"test{k}.txt")
        out = tmp_path / f"bc{k}.txt"
        r = subprocess.run([exe, src, "--bc-out-exact", str(out)], capture_output=True, text=True,
                           cwd=tmp_path, timeout=120)
        assert r.returncode == 0, r.stderr
        lines = r.stdout.splitlines()
        assert lines[0].startswith(" n_bcc= ")
        assert lines[-1].startswith("GPU total time is ")
        assert any(l.startswith("BCC number #: ") for l in lines)
        bc = np.loadtxt(out)[:, 1]
        mx, _, zeros = relerr(bc, goldens[f"sample_test{k}"]["bc"])
        assert mx < BC_TOL and zeros
        assert (tmp_path / "PostProcessOp.txt").exists()


def _expected_stdout(ref):
    """The reference's stdout for an input, numbers that are timings or
    machine properties blanked as 'T': structure lines from the independent
    restatement (tests/golden/make_goldens.py) and, per block of more than
    two vertices, the lines its run prints (Codes/bcc_bfs.cpp:310;
    Codes/kernels_free.cu:10-11,113 and Codes/begin_gpupgm.cu:69 when the
    block has free vertices; Codes/kernels_active.cu:31,217,221 when it has
    chains -- with fewer than 2000 endpoints there is one endpoint batch and
    no postprocess line, Codes/kernels_active.cu:98-158)."""
    lines = [ref["lines"][0]]
    blocks = {b[0]: b for b in ref["blocks"]}
    for l in ref["lines"][1:-1]:
        i = int(l.split()[3])
        _, nf, ne = blocks[i]
        n = int(l.split()[5])
        lines += [l, "Num max threads T"]
        if nf > 0:
            lines += ["Inside bc_gpu_free ", f" Free vertices count is {nf}",
                      "total time for free vertices is ==================== T", "Free vertices processing finished"]
        if ne > 0:
            assert ne < 2000
            lines += [f"max_threads_per_block {min(n, 1024)} number_of_SMs T",
                      "total time for active vertices is T", f"total_act_vert_size is {ne}"]
    return lines + [ref["lines"][-1], "GPU total time is T "]


def _blank(stdout):
    import re
    out = []
    for l in stdout.splitlines():
        if l.startswith("Num max threads ") or l.startswith("max_threads_per_block "):
            l = re.sub(r"\d+$", "T", l)
        elif "time" in l:
            l = re.sub(r"[-+0-9.eE]+(\s*)$", r"T\1", l)
        out.append(l)
    return out


@pytest.mark.parametrize("mode", [[], ["--batched", "--fp32"]])
def test_cli_stdout_matches_reference_restatement(goldens, tmp_path, mode):
    """betc's stdout, line for line, against the reference's stdout restated
    independently: ' n_bcc=', one 'BCC number #:' line per block of more
    than two vertices in Tarjan emission order with the arc count, the
    per-block lines, the active / free totals, timings blanked.  Inputs: the
    four sample inputs and every other golden graph (multi-block graphs with
    bridges, chains and pendants among them), written in file order."""
    exe = os.path.join(bcx.HERE, "betc")
    for name, g in goldens.items():
        src = tmp_path / f"{name}.txt"
        _write_edges(src, g["n"], [tuple(e) for e in g["edges"]])
        out = tmp_path / f"{name}.bc"
        r = subprocess.run([exe, str(src), "--bc-out-exact", str(out)] + mode, capture_output=True, text=True,
                           cwd=tmp_path, timeout=120)
        assert r.returncode == 0, (name, r.stderr)
        assert _blank(r.stdout) == _expected_stdout(g["reference_structure"]), name
        mx, _, zeros = relerr(np.loadtxt(out, ndmin=2)[:, 1], g["bc"])
        assert mx < (1e-6 if mode else BC_TOL) and zeros, name


def test_cli_batched_fp32(goldens, tmp_path):
    exe = os.path.join(bcx.HERE, "betc")
    src = os.path.join(bcx.HERE, "..", "tests", "golden", "test4.txt")
    out = tmp_path / "bc.txt"
    r = subprocess.run([exe, src, "--batched", "--fp32", "--bc-out-exact", str(out)], capture_output=True,
                       text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    mx, _, zeros = relerr(np.loadtxt(out)[:, 1], goldens["sample_test4"]["bc"])
    assert mx < 1e-6 and zeros


def _write_edges(path, n, uv):
    with open(path, "w") as f:
        f.write(f"{n} {len(uv)}\n")
        for a, b in uv:
            f.write(f"{a} {b} 1\n")


def _structure(stdout):
    """stdout with every number after a time label blanked (timings differ)."""
    import re
    out = []
    for l in stdout.splitlines():
        if "time" in l:
            l = re.sub(r"[-+0-9.eE]+\s*$", "T", l)
        out.append(l)
    return out


def test_cli_batched_many_components(tmp_path):
    """The batched CLI uploads every component once and runs them in one
    call: same stdout structure as the per-component plain run (timing values
    aside) and whole-graph BC equal to the oracle's."""
    n, uv = bcx.generate(3, 0, 0.001)
    uv = [tuple(map(int, e)) for e in np.asarray(uv).reshape(-1, 2)]
    src = tmp_path / "g.txt"
    _write_edges(src, n, uv)
    exe = os.path.join(bcx.HERE, "betc")
    rp, col = oracle.csr_from_edges(n, uv)
    want = oracle.brandes_bc(rp, col, threads=16)
    outs = {}
    for name, extra, tol in (("plain", [], 1e-9), ("batched", ["--batched", "--fp32"], 1e-6)):
        out = tmp_path / f"bc_{name}.txt"
        r = subprocess.run([exe, str(src), "--bc-out-exact", str(out)] + extra, capture_output=True, text=True,
                           cwd=tmp_path, timeout=300)
        assert r.returncode == 0, r.stderr
        mx, _, zeros = relerr(np.loadtxt(out)[:, 1], want)
        print(name, "maxrel", mx, flush=True)
        assert mx < tol and zeros
        outs[name] = _structure(r.stdout)
    assert sum(l.startswith("BCC number #: ") for l in outs["plain"]) > 20
    assert outs["plain"] == outs["batched"]


def test_cli_sampled(tmp_path):
    """--sample K: exact when K covers every component; with K below the
    giant component's size, the cut terms stay exact and the sampled,
    rescaled dependency sums stay close to the exact ones."""
    n, uv = bcx.generate(2, 0, 0.003)
    uv = [tuple(map(int, e)) for e in np.asarray(uv).reshape(-1, 2)]
    src = tmp_path / "g.txt"
    _write_edges(src, n, uv)
    exe = os.path.join(bcx.HERE, "betc")
    rp, col = oracle.csr_from_edges(n, uv)
    want = oracle.brandes_bc(rp, col, threads=16)
    h = bcx.Host.from_edges(n, uv).decompose()
    big = max(c.n for c in h.components())
    out = tmp_path / "bc.txt"
    r = subprocess.run([exe, str(src), "--batched", "--sample", str(big), "--bc-out-exact", str(out)],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    mx, _, zeros = relerr(np.loadtxt(out)[:, 1], want)
    assert mx < 1e-9 and zeros
    r = subprocess.run([exe, str(src), "--sample", str(big // 3), "--seed", "7", "--bc-out-exact", str(out)],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    got = np.loadtxt(out)[:, 1]
    assert abs(got.sum() - want.sum()) / want.sum() < 0.05
    top = np.argsort(-want)[:20]
    assert np.all(np.abs(got[top] - want[top]) / want[top] < 0.5)


def test_cli_structure_cache(tmp_path):
    """--cache: the first run writes <input>.bcx, the second reads it; both
    print the same structure lines and the same BC (equal to the oracle's);
    a changed input file is read again."""
    n, uv = bcx.generate(3, 0, 0.001)
    uv = [tuple(map(int, e)) for e in np.asarray(uv).reshape(-1, 2)]
    src = tmp_path / "g.txt"
    _write_edges(src, n, uv)
    exe = os.path.join(bcx.HERE, "betc")
    rp, col = oracle.csr_from_edges(n, uv)
    want = oracle.brandes_bc(rp, col, threads=16)
    runs = []
    for k in range(3):
        if k == 2:
            st = src.stat()
            os.utime(src, ns=(st.st_atime_ns, st.st_mtime_ns + 1_000_000_000))
        out = tmp_path / f"bc{k}.txt"
        r = subprocess.run([exe, str(src), "--batched", "--cache", "--bc-out-exact", str(out)],
                           capture_output=True, text=True, cwd=tmp_path, timeout=300)
        assert r.returncode == 0, r.stderr
        print("run", k, r.stderr.strip(), flush=True)
        mx, _, zeros = relerr(np.loadtxt(out)[:, 1], want)
        assert mx < 1e-9 and zeros
        runs.append((r.stderr, _structure(r.stdout), np.loadtxt(out)[:, 1]))
    assert "structure written to" in runs[0][0]
    assert "structure read from" in runs[1][0]
    assert "structure written to" in runs[2][0]  # stale cache rebuilt
    assert runs[0][1] == runs[1][1] == runs[2][1]
    np.testing.assert_allclose(runs[1][2], runs[0][2], rtol=1e-12, atol=0)  # FP64 sums, tile order may differ
    assert (tmp_path / "g.txt.bcx").exists()
