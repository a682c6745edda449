"""ctypes loader of the CPU oracle (oracle/bcx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the reference result, never as the product.
See the header of bcx_oracle.c for what each function restates and how it is
pinned.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbcx_oracle.so")

_i32p = C.POINTER(C.c_int32)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C {HERE}`")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_source.restype = C.c_int
        L.orc_source.argtypes = [C.c_int32, _i32p, _i32p, _i32p, C.c_int32, _i32p,
                                 C.POINTER(C.c_uint64), dp, dp]
        L.orc_sources.restype = None
        L.orc_sources.argtypes = [C.c_int32, _i32p, _i32p, _i32p, _i32p, C.c_int32, _i32p,
                                  C.POINTER(C.c_uint64), dp, dp, _i32p, C.c_int32]
        L.orc_chain_source.restype = C.c_int
        L.orc_chain_source.argtypes = [C.c_int32, _i32p, _i32p, _i32p, C.c_int32, C.c_int32,
                                       C.c_int32, _i32p, C.c_int32, _i32p,
                                       C.POINTER(C.c_uint64), dp, dp]
        L.orc_component_bc.restype = None
        L.orc_component_bc.argtypes = [C.c_int32, _i32p, _i32p, _i32p, _i32p, C.c_int64, dp,
                                       C.c_int32]
        L.orc_brandes_bc.restype = None
        L.orc_brandes_bc.argtypes = [C.c_int32, _i32p, _i32p, dp, C.c_int32]
        L.orc_cut_term_brute.restype = None
        L.orc_cut_term_brute.argtypes = [C.c_int32, _i32p, _i32p, C.POINTER(C.c_int64)]
        L.orc_reach_brute.restype = None
        L.orc_reach_brute.argtypes = [C.c_int32, _i32p, _i32p, C.POINTER(C.c_uint8), _i32p]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t=C.c_int32):
    return a.ctypes.data_as(C.POINTER(t))


def csr_from_edges(n, uv):
    """Undirected CSR with neighbours in edge-list order (as Codes/graph.cpp:52-57)."""
    uv = np.asarray(uv, np.int64).reshape(-1, 2)
    deg = np.zeros(n + 1, np.int64)
    np.add.at(deg, uv[:, 0] + 1, 1)
    np.add.at(deg, uv[:, 1] + 1, 1)
    rp = np.cumsum(deg).astype(np.int32)
    col = np.zeros(2 * len(uv), np.int32)
    fill = rp[:-1].astype(np.int64).copy()
    for u, v in uv:
        col[fill[u]] = v
        fill[u] += 1
        col[fill[v]] = u
        fill[v] += 1
    return rp, col


def source(rp, col, reach, s):
    """levels, sigma_u64, sigma_f64, delta, overflow of one source."""
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    reach = np.zeros(n, np.int32) if reach is None else np.ascontiguousarray(reach, np.int32)
    lv = np.zeros(n, np.int32)
    su = np.zeros(n, np.uint64)
    sf = np.zeros(n, np.float64)
    dl = np.zeros(n, np.float64)
    ovf = lib().orc_source(n, _p(rp), _p(col), _p(reach), s, _p(lv), _p(su, C.c_uint64),
                           _p(sf, C.c_double), _p(dl, C.c_double))
    return lv, su, sf, dl, bool(ovf)


def sources(rp, col, reach, srcs, threads=0):
    """source() for every s in srcs on host threads: levels [ns, n] int32,
    sigma_u64 [ns, n], sigma_f64 [ns, n], delta [ns, n], overflow [ns] bool."""
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    reach = np.zeros(n, np.int32) if reach is None else np.ascontiguousarray(reach, np.int32)
    srcs = np.ascontiguousarray(srcs, np.int32)
    ns = len(srcs)
    lv = np.zeros((ns, n), np.int32)
    su = np.zeros((ns, n), np.uint64)
    sf = np.zeros((ns, n), np.float64)
    dl = np.zeros((ns, n), np.float64)
    ovf = np.zeros(ns, np.int32)
    lib().orc_sources(n, _p(rp), _p(col), _p(reach), _p(srcs), ns, _p(lv), _p(su, C.c_uint64),
                      _p(sf, C.c_double), _p(dl, C.c_double), _p(ovf), threads)
    return lv, su, sf, dl, ovf.astype(bool)


def chain_source(rp, col, reach, L, R, nodes, pos):
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    reach = np.zeros(n, np.int32) if reach is None else np.ascontiguousarray(reach, np.int32)
    nodes = np.ascontiguousarray(nodes, np.int32)
    lv = np.zeros(n, np.int32)
    su = np.zeros(n, np.uint64)
    sf = np.zeros(n, np.float64)
    dl = np.zeros(n, np.float64)
    ovf = lib().orc_chain_source(n, _p(rp), _p(col), _p(reach), L, R, len(nodes), _p(nodes), pos,
                                 _p(lv), _p(su, C.c_uint64), _p(sf, C.c_double), _p(dl, C.c_double))
    return lv, su, sf, dl, bool(ovf)


def component_bc(rp, col, reach=None, sources=None, threads=0):
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    rptr = None if reach is None else _p(np.ascontiguousarray(reach, np.int32))
    if reach is not None:
        reach = np.ascontiguousarray(reach, np.int32)
        rptr = _p(reach)
    bc = np.zeros(n, np.float64)
    if sources is None:
        lib().orc_component_bc(n, _p(rp), _p(col), rptr, None, 0, _p(bc, C.c_double), threads)
    else:
        src = np.ascontiguousarray(sources, np.int32)
        lib().orc_component_bc(n, _p(rp), _p(col), rptr, _p(src), len(src), _p(bc, C.c_double),
                               threads)
    return bc


def brandes_bc(rp, col, threads=0):
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    bc = np.zeros(n, np.float64)
    lib().orc_brandes_bc(n, _p(rp), _p(col), _p(bc, C.c_double), threads)
    return bc


def cut_term_brute(rp, col):
    n = len(rp) - 1
    out = np.zeros(n, np.int64)
    lib().orc_cut_term_brute(n, _p(np.ascontiguousarray(rp, np.int32)),
                             _p(np.ascontiguousarray(col, np.int32)), _p(out, C.c_int64))
    return out


def reach_brute(rp, col, in_block):
    n = len(rp) - 1
    out = np.zeros(n, np.int32)
    flags = np.ascontiguousarray(in_block, np.uint8)
    lib().orc_reach_brute(n, _p(np.ascontiguousarray(rp, np.int32)),
                          _p(np.ascontiguousarray(col, np.int32)), _p(flags, C.c_uint8), _p(out))
    return out


def num_threads():
    return lib().orc_num_threads()


def pipeline_bc(host, threads=0):
    """Whole-graph BC through the decomposition, all on the CPU: cut term plus
    the weighted BC of every component scattered to global ids (the method of
    Codes/to_eargraph_construct.cpp:151-179 without the doubled free BC)."""
    bc = host.cut_term().astype(np.float64)
    for c in host.components():
        loc = component_bc(c.row_ptr, c.col, c.reach, threads=threads)
        np.add.at(bc, c.gid, loc)
    return bc
