"""Python host mirror of the bcx C-ABI (include/bcx.h), via ctypes.

The library is built in-tree (``make -C centrality-gpu_amd`` or
``__graft_entry__.build()``); loading fails loudly when it is missing, there
is no CPU fallback.  numpy arrays carry the CSR data across the boundary.

Mirrors the reference's operator surface:
  * ``Host.from_file`` / ``Host.from_edges``  -> Graph::Graph (Codes/graph.cpp:38-72)
  * ``Host.decompose``                        -> BiCC + Modified_APSP structure
                                                 (Codes/bicc.cpp:39-161,
                                                  Codes/modified_apsp.cpp:232-373)
  * ``Engine.bc`` / ``Engine.host_bc``        -> begin_gpu free/active + postprocess
                                                 (Codes/begin_gpupgm.cu:42-89)
  * ``Engine.debug_source``                   -> per-source levels/sigma/delta probe
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbcx.so")

BCX_OK = 0
ERRORS = {-1: "BCX_E_ARG", -2: "BCX_E_IO", -3: "BCX_E_GRAPH", -4: "BCX_E_NOMEM",
          -5: "BCX_E_HIP", -6: "BCX_E_NODEV", -7: "BCX_E_OVERFLOW"}
PATH_PLAIN, PATH_BATCHED = 0, 1
PREC_FP64, PREC_FP32C = 0, 1

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class BcxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Component(C.Structure):
    _fields_ = [("n", C.c_int32), ("arcs", C.c_int64), ("row_ptr", _i32p), ("col", _i32p),
                ("reach", _i32p), ("n_chains", C.c_int32), ("chain_ptr", _i32p),
                ("chain_nodes", _i32p), ("chain_ends", _i32p)]


class Opts(C.Structure):
    _fields_ = [("path", C.c_int32), ("precision", C.c_int32), ("use_chains", C.c_int32),
                ("batch_width", C.c_int32), ("src_subset", _i64p), ("n_src", C.c_int64),
                ("row_cap_bytes", C.c_int64), ("sigma_f64_threshold_log2", C.c_int32),
                ("flags", C.c_int32), ("blocks_per_cu", C.c_int32), ("sources_per_lane", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("t_kernel_s", C.c_double), ("t_total_s", C.c_double),
                ("n_bfs_sources", C.c_int64), ("n_chain_sources", C.c_int64),
                ("n_row_only", C.c_int64), ("n_f64_sigma", C.c_int64), ("n_launches", C.c_int64),
                ("max_level", C.c_int32), ("n_ball_tiles", C.c_int32), ("bytes_alg", C.c_double),
                ("teps_work", C.c_double), ("t_chain_s", C.c_double), ("n_graph_launches", C.c_int64),
                ("n_contracted", C.c_int64), ("n_list_entries", C.c_int64), ("n_tiles", C.c_int64),
                ("n_comp_small", C.c_int32), ("n_comp_wg", C.c_int32), ("n_comp_multi", C.c_int32),
                ("n_tuned", C.c_int32), ("sched_blocks_per_cu", C.c_int32), ("sched_slots", C.c_int32),
                ("sched_sources_per_lane", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every symbol include/bcx.h declares, with its ctypes signature
SIGNATURES = {
    "bcx_host_from_file": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "bcx_host_from_edges": (C.c_int, [C.c_int32, C.c_int64, _i32p, C.POINTER(C.c_void_p)]),
    "bcx_host_free": (None, [C.c_void_p]),
    "bcx_host_last_error": (C.c_char_p, [C.c_void_p]),
    "bcx_host_num_vertices": (C.c_int32, [C.c_void_p]),
    "bcx_host_num_edges": (C.c_int64, [C.c_void_p]),
    "bcx_host_decompose": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "bcx_host_num_bcc": (C.c_int32, [C.c_void_p]),
    "bcx_host_bcc_size": (C.c_int, [C.c_void_p, C.c_int32, _i32p, _i64p]),
    "bcx_host_bcc_edges": (C.c_int, [C.c_void_p, C.c_int32, _i32p]),
    "bcx_host_articulation": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint8)]),
    "bcx_host_cut_term": (C.c_int, [C.c_void_p, _i64p]),
    "bcx_host_num_components": (C.c_int32, [C.c_void_p]),
    "bcx_host_component": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Component),
                                     C.POINTER(_i32p), _i32p]),
    "bcx_host_component_counts": (C.c_int, [C.c_void_p, C.c_int32, _i32p, _i32p, _i32p]),
    "bcx_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_size_t]),
    "bcx_destroy": (None, [C.c_void_p]),
    "bcx_last_error": (C.c_char_p, [C.c_void_p]),
    "bcx_device_info": (C.c_int, [C.c_void_p, _i32p, C.c_char_p, C.c_int32]),
    "bcx_upload": (C.c_int, [C.c_void_p, C.POINTER(Component), C.c_int32, C.POINTER(C.c_void_p)]),
    "bcx_dgraph_free": (None, [C.c_void_p, C.c_void_p]),
    "bcx_run": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(C.c_double),
                          C.POINTER(Stats)]),
    "bcx_bc": (C.c_int, [C.c_void_p, C.POINTER(Component), C.c_int32, C.POINTER(Opts),
                         C.POINTER(C.c_double), C.POINTER(Stats)]),
    "bcx_host_bc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(C.c_double),
                              C.POINTER(Stats)]),
    "bcx_debug_source": (C.c_int, [C.c_void_p, C.POINTER(Component), C.c_int32, C.POINTER(Opts),
                                   _i32p, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), _i32p]),
    "bcx_debug_tile": (C.c_int, [C.c_void_p, C.POINTER(Component), _i32p, C.c_int32, C.POINTER(Opts),
                                 _i32p, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), _i32p, _i32p]),
    "bcx_bc_small_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "bcx_generate": (C.c_int, [C.c_int32, C.c_uint64, C.c_double, _i32p, _i64p,
                               C.POINTER(_i32p)]),
    "bcx_batched_order": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, _i32p]),
    "bcx_host_save": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "bcx_host_load": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]),
    "bcx_free": (None, [C.c_void_p]),
}

_lib = None


def lib():
    """Loads libbcx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` "
                               "(or __graft_entry__.build()); there is no fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, t=C.c_int32):
    return a.ctypes.data_as(C.POINTER(t))


def generate(config, seed=0, scale=1.0):
    """(n, uv[m,2] int32) of benchmark workload `config` (1..5)."""
    n = C.c_int32()
    m = C.c_int64()
    p = _i32p()
    rc = lib().bcx_generate(config, seed, scale, C.byref(n), C.byref(m), C.byref(p))
    if rc:
        raise BcxError(rc, "generate")
    uv = np.ctypeslib.as_array(p, shape=(m.value * 2,)).copy().reshape(-1, 2) if m.value else \
        np.zeros((0, 2), np.int32)
    lib().bcx_free(p)
    return n.value, uv


class HostComponent:
    """numpy view of one component (local ids) plus its global ids."""

    def __init__(self, comp, gid, bcc):
        n, a = comp.n, comp.arcs
        self.n, self.arcs, self.bcc = n, a, bcc
        self.row_ptr = np.ctypeslib.as_array(comp.row_ptr, shape=(n + 1,)).copy()
        self.col = np.ctypeslib.as_array(comp.col, shape=(a,)).copy() if a else np.zeros(0, np.int32)
        self.reach = np.ctypeslib.as_array(comp.reach, shape=(n,)).copy()
        k = comp.n_chains
        self.n_chains = k
        self.chain_ptr = np.ctypeslib.as_array(comp.chain_ptr, shape=(k + 1,)).copy()
        tot = int(self.chain_ptr[-1])
        self.chain_nodes = (np.ctypeslib.as_array(comp.chain_nodes, shape=(tot,)).copy()
                            if tot else np.zeros(0, np.int32))
        self.chain_ends = (np.ctypeslib.as_array(comp.chain_ends, shape=(2 * k,)).copy()
                           if k else np.zeros(0, np.int32))
        self.gid = np.ctypeslib.as_array(gid, shape=(n,)).copy()

    def struct(self):
        """ctypes bcx_component pointing at this object's arrays."""
        c = Component()
        c.n, c.arcs = self.n, self.arcs
        self._keep = [np.ascontiguousarray(x, dtype=np.int32) for x in
                      (self.row_ptr, self.col, self.reach, self.chain_ptr, self.chain_nodes,
                       self.chain_ends)]
        rp, col, reach, cp, cn, ce = self._keep
        c.row_ptr, c.col, c.reach = _ptr(rp), _ptr(col), _ptr(reach)
        c.n_chains = self.n_chains
        c.chain_ptr, c.chain_nodes, c.chain_ends = _ptr(cp), _ptr(cn), _ptr(ce)
        return c

    @staticmethod
    def from_csr(row_ptr, col, reach=None, chain_ptr=None, chain_nodes=None, chain_ends=None):
        self = HostComponent.__new__(HostComponent)
        self.row_ptr = np.asarray(row_ptr, np.int32)
        self.col = np.asarray(col, np.int32)
        self.n = len(self.row_ptr) - 1
        self.arcs = len(self.col)
        self.reach = np.zeros(self.n, np.int32) if reach is None else np.asarray(reach, np.int32)
        self.chain_ptr = np.zeros(1, np.int32) if chain_ptr is None else np.asarray(chain_ptr, np.int32)
        self.n_chains = len(self.chain_ptr) - 1
        self.chain_nodes = np.zeros(0, np.int32) if chain_nodes is None else np.asarray(chain_nodes, np.int32)
        self.chain_ends = np.zeros(0, np.int32) if chain_ends is None else np.asarray(chain_ends, np.int32)
        self.gid = np.arange(self.n, dtype=np.int32)
        self.bcc = -1
        return self


class Host:
    """Input graph plus its structural reduction (host C++, no GPU)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)

    @staticmethod
    def from_file(path):
        h = C.c_void_p()
        rc = lib().bcx_host_from_file(path.encode(), C.byref(h))
        if rc:
            raise BcxError(rc, f"cannot read {path}")
        return Host(h.value)

    @staticmethod
    def from_edges(n, uv):
        uv = np.ascontiguousarray(np.asarray(uv, np.int32).reshape(-1))
        h = C.c_void_p()
        rc = lib().bcx_host_from_edges(n, len(uv) // 2, _ptr(uv), C.byref(h))
        if rc:
            raise BcxError(rc, "bad edge list")
        return Host(h.value)

    def save(self, path, source=None):
        """Writes the reduced structure to a cache file (see bcx_host_save)."""
        rc = lib().bcx_host_save(self.h, str(path).encode(), None if source is None else str(source).encode())
        if rc:
            raise BcxError(rc, f"cannot write structure cache {path}")
        return self

    @staticmethod
    def load(path, source=None, use_chains=True):
        """A reduced host from a cache file; BcxError when it is damaged,
        stale for `source` or written with another chains setting."""
        h = C.c_void_p()
        rc = lib().bcx_host_load(str(path).encode(), None if source is None else str(source).encode(),
                                 int(use_chains), C.byref(h))
        if rc:
            raise BcxError(rc, f"structure cache {path} is not usable")
        return Host(h.value)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.bcx_host_free(self.h)
            self.h = C.c_void_p()

    @property
    def n(self):
        return lib().bcx_host_num_vertices(self.h)

    @property
    def m(self):
        return lib().bcx_host_num_edges(self.h)

    def decompose(self, use_chains=True, check_input=True):
        rc = lib().bcx_host_decompose(self.h, int(use_chains), int(check_input))
        if rc:
            raise BcxError(rc, lib().bcx_host_last_error(self.h).decode())
        return self

    def num_bcc(self):
        return lib().bcx_host_num_bcc(self.h)

    def bcc_size(self, i):
        """(vertices, undirected edges) of BCC i."""
        n = C.c_int32()
        e = C.c_int64()
        rc = lib().bcx_host_bcc_size(self.h, i, C.byref(n), C.byref(e))
        if rc:
            raise BcxError(rc, f"no BCC {i}")
        return n.value, e.value

    def num_components(self):
        return lib().bcx_host_num_components(self.h)

    def bcc_edges(self, i):
        n = C.c_int32()
        e = C.c_int64()
        rc = lib().bcx_host_bcc_size(self.h, i, C.byref(n), C.byref(e))
        if rc:
            raise BcxError(rc, f"no BCC {i}")
        uv = np.zeros(2 * e.value, np.int32)
        rc = lib().bcx_host_bcc_edges(self.h, i, _ptr(uv))
        if rc:
            raise BcxError(rc, "BCC edge lists are not available (host loaded from a structure cache?)")
        return uv.reshape(-1, 2)

    def articulation(self):
        f = np.zeros(self.n, np.uint8)
        lib().bcx_host_articulation(self.h, _ptr(f, C.c_uint8))
        return f.astype(bool)

    def cut_term(self):
        c = np.zeros(self.n, np.int64)
        lib().bcx_host_cut_term(self.h, _ptr(c, C.c_int64))
        return c

    def components(self):
        out = []
        for c in range(lib().bcx_host_num_components(self.h)):
            comp = Component()
            gid = _i32p()
            bcc = C.c_int32()
            lib().bcx_host_component(self.h, c, C.byref(comp), C.byref(gid), C.byref(bcc))
            out.append(HostComponent(comp, gid, bcc.value))
        return out

    def component_counts(self, c):
        f, e, x = C.c_int32(), C.c_int32(), C.c_int32()
        lib().bcx_host_component_counts(self.h, c, C.byref(f), C.byref(e), C.byref(x))
        return f.value, e.value, x.value


FLAG_NO_SMALL_KERNEL = 1
FLAG_NO_GRAPH = 2
FLAG_NO_WG_KERNEL = 4
FLAG_WG_KERNEL = 8
FLAG_NO_CONTRACT = 16
FLAG_NO_TUNE = 32
KERNEL_SMALL, KERNEL_WG, KERNEL_MULTI = 1, 2, 3


def make_opts(path=PATH_PLAIN, precision=PREC_FP64, use_chains=True, subset=None,
              row_cap_bytes=0, sigma_threshold_log2=0, batch_width=0, flags=0, blocks_per_cu=0,
              sources_per_lane=0):
    o = Opts()
    o.path, o.precision, o.use_chains = path, precision, int(use_chains)
    o.batch_width = batch_width
    o.flags = flags
    o.blocks_per_cu = blocks_per_cu
    o.sources_per_lane = sources_per_lane
    o.row_cap_bytes = row_cap_bytes
    o.sigma_f64_threshold_log2 = sigma_threshold_log2
    keep = None
    if subset is not None:
        keep = np.ascontiguousarray(np.asarray(subset, np.int64))
        o.src_subset = _ptr(keep, C.c_int64)
        o.n_src = len(keep)
    return o, keep


class Engine:
    """GPU context (device memory, stream, kernels)."""

    def __init__(self, device=0, mem_cap=0):
        h = C.c_void_p()
        rc = lib().bcx_create(C.byref(h), device, mem_cap)
        if rc:
            raise BcxError(rc, "bcx_create failed (no GPU?)")
        self.h = h

    def close(self):
        if self.h and self.h.value:
            lib().bcx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if _lib is not None:
            self.close()

    def _err(self, rc, what):
        raise BcxError(rc, f"{what}: {lib().bcx_last_error(self.h).decode()}")

    def info(self):
        n = C.c_int32()
        name = C.create_string_buffer(128)
        lib().bcx_device_info(self.h, C.byref(n), name, 128)
        return n.value, name.value.decode()

    def upload(self, comps):
        arr = (Component * len(comps))(*[c.struct() for c in comps])
        g = C.c_void_p()
        rc = lib().bcx_upload(self.h, arr, len(comps), C.byref(g))
        if rc:
            self._err(rc, "upload")
        return DGraph(self, g, sum(c.n for c in comps), [c.n for c in comps])

    def bc(self, comps, **kw):
        """Per-component local BC arrays (list) and stats."""
        g = self.upload(comps)
        try:
            flat, st = g.run(**kw)
        finally:
            g.free()
        out, off = [], 0
        for c in comps:
            out.append(flat[off:off + c.n])
            off += c.n
        return out, st

    def host_bc(self, host, **kw):
        o, keep = make_opts(**kw)
        bc = np.zeros(host.n, np.float64)
        st = Stats()
        rc = lib().bcx_host_bc(self.h, host.h, C.byref(o), _ptr(bc, C.c_double), C.byref(st))
        if rc:
            self._err(rc, "host_bc")
        return bc, st.as_dict()

    def debug_source(self, comp, src, **kw):
        o, keep = make_opts(**kw)
        n = comp.n
        lv = np.zeros(n, np.int32)
        su = np.zeros(n, np.uint64)
        sf = np.zeros(n, np.float64)
        dl = np.zeros(n, np.float64)
        isf = C.c_int32()
        cs = comp.struct()
        rc = lib().bcx_debug_source(self.h, C.byref(cs), src, C.byref(o), _ptr(lv),
                                    _ptr(su, C.c_uint64), _ptr(sf, C.c_double),
                                    _ptr(dl, C.c_double), C.byref(isf))
        if rc:
            self._err(rc, "debug_source")
        return lv, su, sf, dl, bool(isf.value)

    def debug_tile(self, comp, srcs, **kw):
        """Every lane of one batched tile (bcx_debug_tile): levels [ns, n]
        int32, sigma_u64 [ns, n], sigma_f64 [ns, n], delta [ns, n], whether
        the counts are FP64, and the BCX_KERNEL_* that ran."""
        kw.setdefault("path", PATH_BATCHED)
        o, keep = make_opts(**kw)
        n = comp.n
        srcs = np.ascontiguousarray(np.asarray(srcs, np.int32))
        ns = len(srcs)
        lv = np.zeros((ns, n), np.int32)
        su = np.zeros((ns, n), np.uint64)
        sf = np.zeros((ns, n), np.float64)
        dl = np.zeros((ns, n), np.float64)
        isf = C.c_int32()
        kern = C.c_int32()
        cs = comp.struct()
        rc = lib().bcx_debug_tile(self.h, C.byref(cs), _ptr(srcs), ns, C.byref(o), _ptr(lv),
                                  _ptr(su, C.c_uint64), _ptr(sf, C.c_double), _ptr(dl, C.c_double),
                                  C.byref(isf), C.byref(kern))
        if rc:
            self._err(rc, "debug_tile")
        return lv, su, sf, dl, bool(isf.value), kern.value


class DGraph:
    def __init__(self, eng, g, total, sizes):
        self.eng, self.g, self.total, self.sizes = eng, g, total, sizes
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)

    def run(self, **kw):
        o, keep = make_opts(**kw)
        bc = np.zeros(max(self.total, 1), np.float64)
        st = Stats()
        rc = lib().bcx_run(self.eng.h, self.g, C.byref(o), _ptr(bc, C.c_double), C.byref(st))
        if rc:
            self.eng._err(rc, "run")
        return bc[:self.total], st.as_dict()

    def batched_order(self, c):
        """Local vertices of component c in the batched path's tile order."""
        out = np.zeros(int(self.sizes[c]), np.int32)
        rc = lib().bcx_batched_order(self.eng.h, self.g, int(c), _ptr(out))
        if rc:
            self.eng._err(rc, "batched_order")
        return out

    def free(self):
        if self.g and self.g.value:
            lib().bcx_dgraph_free(self.eng.h, self.g)
            self.g = C.c_void_p()
