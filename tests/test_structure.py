"""Host structure (include/bcx.h part 1) against networkx goldens, brute
force and invariants; whole-graph parity of the decomposition on the CPU."""
import numpy as np
import pytest

import bcx
import oracle
from helpers import relerr


def test_bcc_and_articulation_vs_networkx(goldens):
    for name, g in goldens.items():
        h = bcx.Host.from_edges(g["n"], g["edges"]).decompose()
        got = sorted(sorted(sorted(map(int, e)) for e in h.bcc_edges(i)) for i in range(h.num_bcc()))
        assert got == g["bccs"], name
        assert list(np.nonzero(h.articulation())[0]) == g["articulation"], name


def test_tarjan_order_sample1():
    # Codes/sample_input/test1.txt is one BCC; the first popped edge is the last
    # back edge pushed by the DFS from vertex 0 in file order.
    h = bcx.Host.from_file(bcx.os.path.join(bcx.HERE, "..", "tests", "golden", "test1.txt")).decompose()
    assert h.num_bcc() == 1
    comps = h.components()
    assert comps[0].n == 9 and comps[0].arcs == 24
    # local ids are first-appearance order over the popped edge list
    uv = h.bcc_edges(0).reshape(-1)
    seen = []
    for v in uv:
        if v not in seen:
            seen.append(int(v))
    assert list(comps[0].gid) == seen


def test_emission_order_local_ids_and_counts_vs_restatement(goldens):
    """Every golden graph (the reference's four sample inputs, multi-block
    graphs with bridges, chains and pendants): BCC sizes in emission order,
    the local ids of every component and the active / free counts equal an
    independent recursive restatement of the reference's Tarjan, relabel and
    ear bookkeeping (tests/golden/make_goldens.py: reference_structure)."""
    multi = 0
    for name, g in goldens.items():
        ref = g["reference_structure"]
        h = bcx.Host.from_edges(g["n"], g["edges"]).decompose()
        assert h.num_bcc() == len(ref["bcc_sizes"]), name
        for i, (nv, ne) in enumerate(ref["bcc_sizes"]):
            assert h.bcc_size(i) == (nv, ne), (name, i)
        big = [i for i, (nv, _) in enumerate(ref["bcc_sizes"]) if nv > 2]
        comps = h.components()
        assert [c.bcc for c in comps] == big, name
        for c in comps:
            assert list(c.gid) == ref["local_ids"][c.bcc], (name, c.bcc)
        active = free = 0
        for k in range(len(comps)):
            f, e, _ = h.component_counts(k)
            active += e
            free += f
        assert ref["lines"][-1] == f"active_vert_count {active} free_vert_count {free} ", name
        multi += len(big) > 1
    assert multi >= 2


def test_cut_term_and_reach_vs_bruteforce(goldens):
    for name, g in goldens.items():
        n = g["n"]
        h = bcx.Host.from_edges(n, g["edges"]).decompose()
        rp, col = oracle.csr_from_edges(n, np.asarray(g["edges"]))
        np.testing.assert_array_equal(h.cut_term(), oracle.cut_term_brute(rp, col), err_msg=name)
        for c in h.components():
            flags = np.zeros(n, np.uint8)
            flags[c.gid] = 1
            want = oracle.reach_brute(rp, col, flags)[c.gid]
            np.testing.assert_array_equal(c.reach, want, err_msg=name)


def chains_by_definition(c):
    deg = np.diff(c.row_ptr)
    if not (deg >= 3).any():
        return set()
    out = set()
    for L in range(c.n):
        if deg[L] < 3:
            continue
        for j in range(c.row_ptr[L], c.row_ptr[L + 1]):
            x = int(c.col[j])
            if deg[x] != 2:
                continue
            path, prev, cur = [], L, x
            while deg[cur] == 2:
                path.append(cur)
                a, b = c.col[c.row_ptr[cur]], c.col[c.row_ptr[cur] + 1]
                prev, cur = cur, int(a if a != prev else b)
            key = tuple(path) if path[0] < path[-1] or len(path) == 1 and L < cur else tuple(path[::-1])
            out.add((frozenset([L, cur]), frozenset(path)))
    return out


@pytest.mark.parametrize("cfg,scale", [(1, 0.25), (2, 0.003), (4, 0.05), (5, 0.004)])
def test_chain_invariants(cfg, scale):
    n, uv = bcx.generate(cfg, 0, scale)
    h = bcx.Host.from_edges(n, uv).decompose()
    for ci, c in enumerate(h.components()):
        got = set()
        deg = np.diff(c.row_ptr)
        for k in range(c.n_chains):
            nodes = c.chain_nodes[c.chain_ptr[k]:c.chain_ptr[k + 1]]
            L, R = int(c.chain_ends[2 * k]), int(c.chain_ends[2 * k + 1])
            assert L != R and deg[L] >= 3 and deg[R] >= 3
            assert (deg[nodes] == 2).all()
            # consecutive along the path, ld = j+1 / rd = cnt-j distances
            seq = [L] + list(map(int, nodes)) + [R]
            for a, b in zip(seq, seq[1:]):
                assert b in c.col[c.row_ptr[a]:c.row_ptr[a + 1]]
            got.add((frozenset([L, R]), frozenset(map(int, nodes))))
        assert got == chains_by_definition(c)
        f, e, x = h.component_counts(ci)
        assert f + e + x == c.n
        assert x == len(c.chain_nodes)


@pytest.mark.parametrize("cfg,scale", [(1, 0.25), (2, 0.002), (4, 0.02), (5, 0.002)])
def test_pipeline_equals_whole_graph_brandes(cfg, scale):
    """cut term + weighted component BC == plain Brandes on the input graph."""
    n, uv = bcx.generate(cfg, 0, scale)
    h = bcx.Host.from_edges(n, uv).decompose()
    rp, col = oracle.csr_from_edges(n, uv)
    whole = oracle.brandes_bc(rp, col, threads=4)
    piped = oracle.pipeline_bc(h, threads=4)
    mx, mean, zeros = relerr(piped, whole)
    assert mx < 1e-12 and zeros


def test_input_checks():
    with pytest.raises(bcx.BcxError):
        bcx.Host.from_edges(3, [(0, 1), (1, 1)]).decompose()
    with pytest.raises(bcx.BcxError):
        bcx.Host.from_edges(3, [(0, 1), (1, 0), (1, 2)]).decompose()
    with pytest.raises(bcx.BcxError):
        bcx.Host.from_edges(4, [(0, 1), (2, 3)]).decompose()
    with pytest.raises(bcx.BcxError):
        bcx.Host.from_file("/nonexistent/graph.txt")


def test_file_reader_matches_edges(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("4 4\n0 1 1\n1 2 1\n2 3 1\n3 0 1\n")
    h = bcx.Host.from_file(str(p)).decompose()
    assert h.n == 4 and h.m == 4 and h.num_bcc() == 1
    p.write_text("4 2\n0 1 1\n1 2 2\n")
    with pytest.raises(bcx.BcxError):
        bcx.Host.from_file(str(p))


def test_input_check_messages():
    # first vertex whose list repeats a neighbour, reported at the repeat
    with pytest.raises(bcx.BcxError, match="duplicate edge 1 2"):
        bcx.Host.from_edges(4, [(0, 1), (1, 2), (2, 0), (2, 1), (2, 3)]).decompose()
    with pytest.raises(bcx.BcxError, match="self loop on vertex 2"):
        bcx.Host.from_edges(4, [(0, 1), (2, 2), (1, 2), (3, 3)]).decompose()
    with pytest.raises(bcx.BcxError, match=r"disconnected \(2 of 5 vertices reachable from 0\)"):
        bcx.Host.from_edges(5, [(0, 1), (2, 3), (3, 4), (4, 2)]).decompose()


def test_file_reader_token_rules(tmp_path):
    """Token-by-token reading (fscanf semantics of Codes/graph.cpp:44-59):
    line breaks are whitespace, anything after the m-th edge is ignored, a
    short or malformed edge list is an error."""
    p = tmp_path / "g.txt"

    def read(text):
        p.write_text(text)
        return bcx.Host.from_file(str(p))

    ok = {
        "plain": "3 3\n0 1 1\n1 2 1\n2 0 1\n",
        "no final newline": "3 3\n0 1 1\n1 2 1\n2 0 1",
        "split lines": "3 3 0\n1 1 1 2\n1\n2 0 1\n",
        "signs and tabs": "+3\t3\r\n0 +1 1\n1 2 +1\n2 0 1\n",
        "trailing junk": "3 3\n0 1 1\n1 2 1\n2 0 1\nthis is not an edge\n",
        "glued last token": "3 3\n0 1 1\n1 2 1\n2 0 1x\n",
    }
    for name, text in ok.items():
        h = read(text).decompose()
        assert (h.n, h.m, h.num_bcc()) == (3, 3, 1), name
        assert sorted(map(tuple, np.sort(h.bcc_edges(0), axis=1).tolist())) == [(0, 1), (0, 2), (1, 2)], name
    bad = {
        "short": "3 3\n0 1 1\n1 2 1\n",
        "letter inside": "3 3\n0 1 1\n1 x 1\n2 0 1\n",
        "id out of range": "3 3\n0 1 1\n1 3 1\n2 0 1\n",
        "negative id": "3 3\n0 1 1\n1 -2 1\n2 0 1\n",
        "weight 2": "3 3\n0 1 1\n1 2 2\n2 0 1\n",
        "empty": "",
        "header only n": "3",
    }
    for name, text in bad.items():
        with pytest.raises(bcx.BcxError):
            read(text)
            pytest.fail(name)


def _structure_arrays(h):
    out = [h.cut_term(), h.articulation()]
    for c in h.components():
        out += [np.asarray(x) for x in (c.gid, c.row_ptr, c.col, c.reach, c.chain_ptr, c.chain_nodes, c.chain_ends)]
    return out


def test_threaded_reader_and_structure_match_single_thread(tmp_path, monkeypatch):
    """The text reader scans whitespace-cut ranges on several threads and the
    adjacency, input check and component build run threaded: every array must
    equal the single-thread result and the in-memory edge list's."""
    for cfg, scale in ((2, 0.03), (3, 0.002)):
        n, uv = bcx.generate(cfg, 0, scale)
        p = tmp_path / f"g{cfg}.txt"
        with open(p, "w") as f:
            f.write(f"{n} {len(uv)}\n")
            np.savetxt(f, np.column_stack([uv, np.ones(len(uv), np.int64)]), fmt="%d")
        assert p.stat().st_size > (1 << 20)  # large enough for the threaded scan
        results = []
        for threads in ("1", "3", "7"):
            monkeypatch.setenv("BCX_HOST_THREADS", threads)
            results.append(_structure_arrays(bcx.Host.from_file(str(p)).decompose()))
            results.append(_structure_arrays(bcx.Host.from_edges(n, uv).decompose()))
        for r in results[1:]:
            assert len(r) == len(results[0])
            for a, b in zip(r, results[0]):
                np.testing.assert_array_equal(a, b)


def _write_text(path, n, uv):
    with open(path, "w") as f:
        f.write(f"{n} {len(uv)}\n")
        np.savetxt(f, np.column_stack([uv, np.ones(len(uv), np.int64)]), fmt="%d")


def test_structure_cache_round_trip(tmp_path):
    """A saved and reloaded host answers every query as the original did."""
    for cfg, scale in ((1, 1.0), (3, 0.002), (4, 0.05)):
        n, uv = bcx.generate(cfg, 0, scale)
        src = tmp_path / f"g{cfg}.txt"
        _write_text(src, n, uv)
        h = bcx.Host.from_file(str(src)).decompose()
        cache = tmp_path / f"g{cfg}.txt.bcx"
        h.save(cache, source=src)
        g = bcx.Host.load(cache, source=src)
        assert (g.n, g.m, g.num_bcc(), g.num_components()) == (h.n, h.m, h.num_bcc(), h.num_components())
        for a, b in zip(_structure_arrays(g), _structure_arrays(h)):
            np.testing.assert_array_equal(a, b)
        for i in range(h.num_bcc()):
            assert g.bcc_size(i) == h.bcc_size(i)
        for c in range(h.num_components()):
            assert g.component_counts(c) == h.component_counts(c)
        g.decompose()  # same chains setting: nothing to do
        with pytest.raises(bcx.BcxError):
            g.bcc_edges(0)  # edge lists are not stored
        if cfg == 1:
            np.testing.assert_array_equal(oracle.pipeline_bc(g, threads=1), oracle.pipeline_bc(h, threads=1))


def test_structure_cache_rejects_stale_or_damaged(tmp_path):
    n, uv = bcx.generate(1, 0, 1.0)
    src = tmp_path / "g.txt"
    _write_text(src, n, uv)
    cache = tmp_path / "g.txt.bcx"
    bcx.Host.from_file(str(src)).decompose().save(cache, source=src)
    bcx.Host.load(cache, source=src)
    bcx.Host.load(cache)  # no source: format checks only
    with pytest.raises(bcx.BcxError):
        bcx.Host.load(cache, source=src, use_chains=False)
    st = src.stat()
    bcx.os.utime(src, ns=(st.st_atime_ns, st.st_mtime_ns + 1_000_000_000))
    with pytest.raises(bcx.BcxError):
        bcx.Host.load(cache, source=src)  # the text changed after the cache was written
    raw = cache.read_bytes()
    for bad in (raw[:-8], raw[: len(raw) // 2], raw + b"\0" * 8, b"BCXSTRUX" + raw[8:], b""):
        damaged = tmp_path / "damaged.bcx"
        damaged.write_bytes(bad)
        with pytest.raises(bcx.BcxError):
            bcx.Host.load(damaged)
    # a column index out of range is caught before any view is handed out
    h = bcx.Host.load(cache)
    col0 = np.asarray(h.components()[0].col)
    arr = np.frombuffer(raw, np.uint8).copy()
    needle = col0[:16].astype("<i4").tobytes()
    at = raw.find(needle)
    assert at > 0
    arr[at:at + 4] = np.frombuffer(np.int32(10**8).astype("<i4").tobytes(), np.uint8)
    damaged.write_bytes(arr.tobytes())
    with pytest.raises(bcx.BcxError):
        bcx.Host.load(damaged)
    with pytest.raises(bcx.BcxError):
        bcx.Host.from_edges(3, [(0, 1), (1, 2), (2, 0)]).save(tmp_path / "x.bcx")  # not decomposed


def test_generators_deterministic_and_simple():
    for cfg in (1, 2, 3, 4, 5):
        n1, uv1 = bcx.generate(cfg, 0, 0.002 if cfg != 1 else 0.2)
        n2, uv2 = bcx.generate(cfg, 0, 0.002 if cfg != 1 else 0.2)
        assert n1 == n2 and np.array_equal(uv1, uv2)
        bcx.Host.from_edges(n1, uv1).decompose(check_input=True)  # simple + connected
