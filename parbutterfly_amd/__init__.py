"""parbutterfly_amd -- butterfly counting and peeling on AMD MI355X (gfx950).

Python view of the C ABI in include/parbutterfly.h (libparbutterfly_amd.so).  Device memory is
the library's own (DeviceArray / DeviceGraph); torch is not needed.  It mirrors the
reference's interface for this path:

  reference (apps/hyper)                       here
  -----------------------------------------    -------------------------------------------
  readBipartite (butterfly_utils.h:1008)       Graph.read(path)
  cmpWedgeCounts (butterfly_utils.h:1080)      Context.side(g) -> (use_v, num_wedges)
  getRanks / getApproxDegRanks / ...           Context.rank(g, rank_type)
  Count (butterfly_count.h:1634)               Context.count(g, use_v, "VERT", ...)
  CountE (butterfly_ecount.h:1766)             Context.count(g, use_v, "EDGE", ...)
  CountTotal (butterfly_count_total.h:1014)    Context.count(g, use_v, "TOTAL", ...)
  edgeToIdxs (butterfly_utils.h:2034)          Context.edge_to_idxs(g, use_v)
  Peel (butterfly_peel.h:896)                  Context.peel(g, use_v, counts)
  PeelE (butterfly_epeel.h:519)                Context.peel_edge(g, use_v, counts)

Count / rank / peel types accept the reference's names ("BATCHS", "ADEG", ...) or its numeric
enum values; every count type runs on the GPU ("GPU" names the mode explicitly).  Errors raise
PbfError carrying the library's status code; nothing falls back to a CPU path: if the shared
library is missing, importing this package raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libparbutterfly_amd.so")

if not os.path.exists(_LIB_PATH):
    raise ImportError(
        f"{_LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `python parbutterfly_amd/build.py` first")
_lib = ctypes.CDLL(_LIB_PATH)

# enums (apps/hyper/butterfly_utils.h:24-28, values kept)
COUNT_TYPES = {"SORT": 0, "ASORT": 1, "HASH": 2, "AHASH": 3, "HIST": 4, "AHIST": 5, "BATCHS": 6,
               "BATCHWA": 7, "SERIAL": 8, "GPU": 9}
RANK_TYPES = {"SIDE": 0, "COCORE": 1, "ACOCORE": 2, "DEG": 3, "ADEG": 4}
PEEL_TYPES = {"SORT": 0, "HASH": 1, "HIST": 2, "BATCHS": 3, "BATCHWA": 4, "GPU": 5}
PER_TYPES = {"VERT": 0, "EDGE": 1, "TOTAL": 2}
SPARSE_TYPES = {"NONE": 0, "COLOR": 1, "EDGE": 2}  # SparseType, butterfly_utils.h:28
MEM_HOST, MEM_DEVICE = 0, 1

STATUS = {0: "PBF_OK", 1: "PBF_ERR_ARG", 2: "PBF_ERR_FORMAT", 3: "PBF_ERR_RANGE", 4: "PBF_ERR_UNSORTED",
          5: "PBF_ERR_OOM", 6: "PBF_ERR_CAPACITY", 7: "PBF_ERR_DOMAIN", 8: "PBF_ERR_HIP",
          9: "PBF_ERR_UNSUPPORTED", 10: "PBF_ERR_IO"}


class PbfError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        self.name = STATUS.get(status, str(status))
        msg = _lib.pbf_status_str(status).decode()
        super().__init__(f"{where}: {self.name} ({msg})")


class _CSR(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int64), ("nu", ctypes.c_int64), ("m", ctypes.c_int64),
                ("offsetsV", ctypes.c_void_p), ("edgesV", ctypes.c_void_p),
                ("offsetsU", ctypes.c_void_p), ("edgesU", ctypes.c_void_p)]


class _Graph(ctypes.Structure):
    _fields_ = [("csr", _CSR), ("storage", ctypes.c_void_p)]


class Timing(ctypes.Structure):
    _fields_ = [("rank_ms", ctypes.c_double), ("relabel_ms", ctypes.c_double), ("count_ms", ctypes.c_double),
                ("convert_ms", ctypes.c_double), ("peel_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("wedges", ctypes.c_int64), ("rounds", ctypes.c_int64), ("hub_vertices", ctypes.c_int64),
                ("sparsify_ms", ctypes.c_double), ("batches", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_INT = ctypes.c_int
class SelectorModel(ctypes.Structure):
    """pbf_selector_model: network parameters and the standardisation of inputs and outputs."""
    _fields_ = [("params", ctypes.c_float * 6288), ("fmu", ctypes.c_float * 16), ("fsd", ctypes.c_float * 16),
                ("tmu", ctypes.c_float * 9), ("tsd", ctypes.c_float * 9)]

    @classmethod
    def from_dict(cls, m):
        c = cls()
        for k in ("params", "fmu", "fsd", "tmu", "tsd"):
            a = np.ascontiguousarray(m[k], np.float32)
            if a.size != len(getattr(c, k)):
                raise ValueError(f"selector model field {k}: {a.size} values")
            ctypes.memmove(getattr(c, k), a.ctypes.data, a.nbytes)
        return c


_PI = ctypes.POINTER(ctypes.c_int)
_PI64 = ctypes.POINTER(ctypes.c_int64)
_PCSR = ctypes.POINTER(_CSR)
_PG = ctypes.POINTER(_Graph)

_SIGS = {
    "pbf_status_str": (ctypes.c_char_p, [_INT]),
    "pbf_abi_version": (_INT, []),
    "pbf_read_bipartite": (_INT, [ctypes.c_char_p, _PG]),
    "pbf_read_konect": (_INT, [ctypes.c_char_p, _PG]),
    "pbf_write_bipartite": (_INT, [ctypes.c_char_p, _PCSR]),
    "pbf_write_binary": (_INT, [ctypes.c_char_p, _PCSR]),
    "pbf_validate": (_INT, [_PCSR]),
    "pbf_graph_alloc": (_INT, [_I64, _I64, _I64, _PG]),
    "pbf_graph_free": (None, [_PG]),
    "pbf_graph_from_edges": (_INT, [_I64, _I64, _I64, _P, _P, _PG]),
    "pbf_gen_uniform": (_INT, [_I64, _I64, _I64, _U64, _PG]),
    "pbf_gen_chung_lu": (_INT, [_I64, _I64, _I64, ctypes.c_double, ctypes.c_double, _U64, _PG]),
    "pbf_gen_rmat": (_INT, [_INT, _INT, _I64, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                            _U64, _PG]),
    "pbf_batch_graph_params": (_INT, [_U64, _PI64, _PI64, _PI64, ctypes.POINTER(ctypes.c_double)]),
    "pbf_gen_batch_graph": (_INT, [_U64, _PG]),
    "pbf_side_host": (_INT, [_PCSR, _PI, _PI64]),
    "pbf_create": (_INT, [ctypes.POINTER(_P), _INT]),
    "pbf_destroy": (_INT, [_P]),
    "pbf_stream": (_P, [_P]),
    "pbf_set_stream": (_INT, [_P, _P]),
    "pbf_synchronize": (_INT, [_P]),
    "pbf_last_timing": (_INT, [_P, ctypes.POINTER(Timing)]),
    "pbf_trim": (_INT, [_P]),
    "pbf_set_limits": (_INT, [_P, _I64, _I64]),
    "pbf_device_alloc": (_INT, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "pbf_device_free": (_INT, [_P, _P]),
    "pbf_copy_to_device": (_INT, [_P, _P, _P, ctypes.c_size_t]),
    "pbf_copy_to_host": (_INT, [_P, _P, _P, ctypes.c_size_t]),
    "pbf_copy_device": (_INT, [_P, _P, _P, ctypes.c_size_t]),
    "pbf_side": (_INT, [_P, _PCSR, _INT, _PI, _PI64]),
    "pbf_rank": (_INT, [_P, _PCSR, _INT, _INT, _P, _P, _PI64]),
    "pbf_count_vert": (_INT, [_P, _PCSR, _INT, _INT, _INT, _INT, _P]),
    "pbf_count_vert_noalloc": (_INT, [_P, _PCSR, _INT, _INT, _INT, _P]),
    "pbf_count_edge": (_INT, [_P, _PCSR, _INT, _INT, _INT, _INT, _P]),
    "pbf_count_total": (_INT, [_P, _PCSR, _INT, _INT, _INT, _INT, _P]),
    "pbf_edge_to_idxs": (_INT, [_P, _PCSR, _INT, _INT, _P]),
    "pbf_peel_vert": (_INT, [_P, _PCSR, _INT, _INT, _INT, _P, _P]),
    "pbf_peel_edge": (_INT, [_P, _PCSR, _INT, _INT, _INT, _P, _P]),
    "pbf_read_bipartite_device": (_INT, [_P, ctypes.c_char_p, _PCSR, ctypes.POINTER(_P)]),
    "pbf_sparsify": (_INT, [_P, _PCSR, _INT, _INT, _I64, _I64, _PI64, _PI64, _PI64, _P, _P, _P, _P]),
    "pbf_count_total_sparse": (_INT, [_P, _PCSR, _INT, _INT, _I64, _I64, _INT, _INT, _PI64,
                                      ctypes.POINTER(ctypes.c_uint64)]),
    "pbf_count_batch": (_INT, [_P, _I64, _PCSR, _INT, _INT, _INT, _P, _I64, _P, _P, _INT]),
    "pbf_batch_release": (None, [_P]),
    "pbf_batch_autotune": (_INT, [_P, _I64, _PCSR, _INT, _INT, _P, _I64, _P, _P, _INT, _PI,
                                  ctypes.POINTER(ctypes.c_float)]),
    "pbf_set_tuning": (_INT, [_P, _INT, _INT, _INT]),
    "pbf_batch_features": (_INT, [_P, _I64, _PCSR, _INT, _P]),
    "pbf_variant_tuning": (_INT, [_INT, _PI, _PI, _PI, _PI]),
    "pbf_mlp_init": (_INT, [_U64, _P]),
    "pbf_mlp_train": (_INT, [_P, _P, _P, _P, _I64, _INT, ctypes.c_float, _U64, ctypes.POINTER(ctypes.c_float)]),
    "pbf_mlp_grad": (_INT, [_P, _P, _P, _P, _I64, _P, ctypes.POINTER(ctypes.c_float)]),
    "pbf_mlp_predict": (_INT, [_P, _P, _P, _I64, _P]),
    "pbf_select": (_INT, [_P, _I64, _PCSR, _INT, ctypes.POINTER(SelectorModel), _P]),
    "pbf_count_batch_learned": (_INT, [_P, _I64, _PCSR, _INT, ctypes.POINTER(SelectorModel), _P, _I64, _P, _P, _INT,
                                       _P]),
}
for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

EXPORTED_SYMBOLS = tuple(_SIGS)


def _check(status, where):
    if status != 0:
        raise PbfError(status, where)


def _enum(table, value, what):
    if isinstance(value, str):
        if value not in table:
            raise ValueError(f"unknown {what} {value!r}")
        return table[value]
    return int(value)


def library_path():
    return _LIB_PATH


class Graph:
    """Host bipartite graph in the reference's bipartiteCSR layout (uint32 offsets / ids).

    Arrays are numpy views.  Graphs made by the library (reader, generators) own C storage;
    graphs made from numpy arrays keep references to them.
    """

    def __init__(self, nv, nu, offsetsV, edgesV, offsetsU, edgesU, _owner=None):
        self.nv, self.nu = int(nv), int(nu)
        self.offsetsV = np.ascontiguousarray(offsetsV, dtype=np.uint32)
        self.edgesV = np.ascontiguousarray(edgesV, dtype=np.uint32)
        self.offsetsU = np.ascontiguousarray(offsetsU, dtype=np.uint32)
        self.edgesU = np.ascontiguousarray(edgesU, dtype=np.uint32)
        self.m = int(self.edgesV.shape[0])
        self._owner = _owner
        self._csr = _CSR(self.nv, self.nu, self.m, self.offsetsV.ctypes.data, self.edgesV.ctypes.data,
                         self.offsetsU.ctypes.data, self.edgesU.ctypes.data)

    # -- construction --------------------------------------------------------------------
    @staticmethod
    def _adopt(g):
        c = g.csr
        nv, nu, m = c.nv, c.nu, c.m

        def arr(ptr, n):
            if n == 0:
                return np.zeros(0, dtype=np.uint32)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint32)), shape=(n,))

        owner = _Owner(g)
        return Graph(nv, nu, arr(c.offsetsV, nv + 1), arr(c.edgesV, m), arr(c.offsetsU, nu + 1), arr(c.edgesU, m),
                     _owner=owner)

    @staticmethod
    def read(path):
        """readBipartite: AdjacencyHypergraph text or PBFCSR1 binary."""
        g = _Graph()
        _check(_lib.pbf_read_bipartite(os.fsencode(path), ctypes.byref(g)), "pbf_read_bipartite")
        return Graph._adopt(g)

    @staticmethod
    def read_konect(path):
        """pbf_read_konect: a KONECT bipartite edge list (readKONECT + KONECTtoHyperAdj)."""
        g = _Graph()
        _check(_lib.pbf_read_konect(os.fsencode(path), ctypes.byref(g)), "pbf_read_konect")
        return Graph._adopt(g)

    @staticmethod
    def from_edges(nv, nu, v, u):
        v = np.ascontiguousarray(v, dtype=np.uint32)
        u = np.ascontiguousarray(u, dtype=np.uint32)
        g = _Graph()
        _check(_lib.pbf_graph_from_edges(nv, nu, v.shape[0], v.ctypes.data, u.ctypes.data, ctypes.byref(g)),
               "pbf_graph_from_edges")
        return Graph._adopt(g)

    @staticmethod
    def uniform(nv, nu, m, seed):
        g = _Graph()
        _check(_lib.pbf_gen_uniform(nv, nu, m, seed, ctypes.byref(g)), "pbf_gen_uniform")
        return Graph._adopt(g)

    @staticmethod
    def chung_lu(nv, nu, m, gamma_v, gamma_u, seed):
        g = _Graph()
        _check(_lib.pbf_gen_chung_lu(nv, nu, m, gamma_v, gamma_u, seed, ctypes.byref(g)), "pbf_gen_chung_lu")
        return Graph._adopt(g)

    @staticmethod
    def rmat(log_nv, log_nu, m, a, b, c, d, seed):
        g = _Graph()
        _check(_lib.pbf_gen_rmat(log_nv, log_nu, m, a, b, c, d, seed, ctypes.byref(g)), "pbf_gen_rmat")
        return Graph._adopt(g)

    @staticmethod
    def batch_params(seed):
        nv, nu, m, gamma = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        _check(_lib.pbf_batch_graph_params(seed, ctypes.byref(nv), ctypes.byref(nu), ctypes.byref(m),
                                           ctypes.byref(gamma)), "pbf_batch_graph_params")
        return nv.value, nu.value, m.value, gamma.value

    @staticmethod
    def batch_graph(seed):
        g = _Graph()
        _check(_lib.pbf_gen_batch_graph(seed, ctypes.byref(g)), "pbf_gen_batch_graph")
        return Graph._adopt(g)

    # -- IO / checks ---------------------------------------------------------------------
    def write(self, path):
        _check(_lib.pbf_write_bipartite(os.fsencode(path), ctypes.byref(self._csr)), "pbf_write_bipartite")

    def write_binary(self, path):
        _check(_lib.pbf_write_binary(os.fsencode(path), ctypes.byref(self._csr)), "pbf_write_binary")

    def validate(self):
        return _lib.pbf_validate(ctypes.byref(self._csr))

    def side(self):
        """cmpWedgeCounts on the host: (use_v, num_wedges)."""
        uv, nw = ctypes.c_int(), ctypes.c_int64()
        _check(_lib.pbf_side_host(ctypes.byref(self._csr), ctypes.byref(uv), ctypes.byref(nw)), "pbf_side_host")
        return bool(uv.value), nw.value

    @property
    def n(self):
        return self.nv + self.nu

    def to_device(self, ctx):
        """Copies the four arrays to HBM on ctx's device; returns a DeviceGraph."""
        return DeviceGraph(self, ctx)


class _Owner:
    def __init__(self, g):
        self.g = g

    def __del__(self):
        try:
            _lib.pbf_graph_free(ctypes.byref(self.g))
        except Exception:
            pass


class DeviceArray:
    """A 1-D array in device memory, allocated with the library's own allocator."""

    def __init__(self, ctx, n, dtype, host=None):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(_lib.pbf_device_alloc(ctx._ctx, max(self.nbytes, 16), ctypes.byref(p)), "pbf_device_alloc")
        self.ptr = p.value
        if host is not None:
            h = np.ascontiguousarray(host, dtype=self.dtype)
            assert h.size == self.n
            if self.nbytes:
                _check(_lib.pbf_copy_to_device(ctx._ctx, self.ptr, h.ctypes.data, self.nbytes), "pbf_copy_to_device")

    @classmethod
    def _view(cls, ctx, ptr, n, dtype, base):
        """n elements at device address ptr inside `base` (a DeviceArray that owns the memory)."""
        self = cls.__new__(cls)
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = ptr
        self._base = base
        return self

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        if self.nbytes:
            _check(_lib.pbf_copy_to_host(self.ctx._ctx, out.ctypes.data, self.ptr, self.nbytes), "pbf_copy_to_host")
        return out

    def copy_from(self, other):
        """Device-to-device copy of another DeviceArray of the same size."""
        assert other.nbytes == self.nbytes
        if self.nbytes:
            _check(_lib.pbf_copy_device(self.ctx._ctx, self.ptr, other.ptr, self.nbytes), "pbf_copy_device")

    def free(self):
        if getattr(self, "_base", None) is not None:  # a view: the base owns the memory
            self.ptr = None
            return
        if self.ptr and self.ctx._ctx:
            _lib.pbf_device_free(self.ctx._ctx, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self):
        return self.n


class DeviceGraph:
    """A graph whose four arrays live in HBM (PBF_MEM_DEVICE)."""

    def __init__(self, g, ctx):
        self.nv, self.nu, self.m = g.nv, g.nu, g.m
        self.ctx = ctx
        self.offsetsV = DeviceArray(ctx, g.nv + 1, np.uint32, g.offsetsV)
        self.edgesV = DeviceArray(ctx, g.m, np.uint32, g.edgesV)
        self.offsetsU = DeviceArray(ctx, g.nu + 1, np.uint32, g.offsetsU)
        self.edgesU = DeviceArray(ctx, g.m, np.uint32, g.edgesU)
        self._csr = _CSR(self.nv, self.nu, self.m, self.offsetsV.ptr, self.edgesV.ptr, self.offsetsU.ptr,
                         self.edgesU.ptr)

    @property
    def n(self):
        return self.nv + self.nu

    @classmethod
    def read(cls, ctx, path):
        """readBipartite straight into HBM (text parsed and everything validated on the GPU)."""
        csr, blk = _CSR(), ctypes.c_void_p()
        _check(_lib.pbf_read_bipartite_device(ctx._ctx, os.fsencode(path), ctypes.byref(csr), ctypes.byref(blk)),
               "pbf_read_bipartite_device")
        base = DeviceArray._view(ctx, blk.value, 0, np.uint8, None)
        base._base = None  # owns the block
        arrs = [DeviceArray._view(ctx, getattr(csr, k), n, np.uint32, base) for k, n in
                (("offsetsV", csr.nv + 1), ("edgesV", csr.m), ("offsetsU", csr.nu + 1), ("edgesU", csr.m))]
        g = cls._from_arrays(ctx, csr.nv, csr.nu, csr.m, *arrs)
        g._block = base
        return g

    @classmethod
    def _from_arrays(cls, ctx, nv, nu, m, offsetsV, edgesV, offsetsU, edgesU):
        """Wraps device arrays (possibly longer than nv + 1 / m / nu + 1 / m) as a graph."""
        self = cls.__new__(cls)
        self.nv, self.nu, self.m, self.ctx = int(nv), int(nu), int(m), ctx
        self.offsetsV, self.edgesV, self.offsetsU, self.edgesU = offsetsV, edgesV, offsetsU, edgesU
        self._csr = _CSR(self.nv, self.nu, self.m, offsetsV.ptr, edgesV.ptr, offsetsU.ptr, edgesU.ptr)
        return self


class Context:
    """A GPU context (device, stream, cached scratch).  One per host thread."""

    def __init__(self, device=0):
        self._ctx = ctypes.c_void_p()
        _check(_lib.pbf_create(ctypes.byref(self._ctx), int(device)), "pbf_create")
        self.device = int(device)

    def close(self):
        if self._ctx:
            _lib.pbf_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self):
        return _lib.pbf_stream(self._ctx)

    def set_stream(self, stream):
        """Queues the context's work on a caller's hipStream_t (an integer handle); None
        returns to the context's own stream."""
        _check(_lib.pbf_set_stream(self._ctx, stream), "pbf_set_stream")

    def synchronize(self):
        _check(_lib.pbf_synchronize(self._ctx), "pbf_synchronize")

    def trim(self):
        _check(_lib.pbf_trim(self._ctx), "pbf_trim")

    def set_limits(self, max_wedges=0, max_array_size=0):
        """pbf_set_limits: the reference's -m / -a bounds of the counting passes (0 = none)."""
        _check(_lib.pbf_set_limits(self._ctx, int(max_wedges), int(max_array_size)), "pbf_set_limits")

    def timing(self):
        t = Timing()
        _check(_lib.pbf_last_timing(self._ctx, ctypes.byref(t)), "pbf_last_timing")
        return t.as_dict()

    @staticmethod
    def _mem(g):
        return MEM_DEVICE if isinstance(g, DeviceGraph) else MEM_HOST

    def _alloc_out(self, g, n, dtype):
        if isinstance(g, DeviceGraph):
            return DeviceArray(self, n, dtype)
        return np.zeros(n, dtype=dtype)

    @staticmethod
    def _ptr(a):
        if isinstance(a, np.ndarray):
            return a.ctypes.data
        return a.ptr

    def side(self, g):
        uv, nw = ctypes.c_int(), ctypes.c_int64()
        _check(_lib.pbf_side(self._ctx, ctypes.byref(g._csr), self._mem(g), ctypes.byref(uv), ctypes.byref(nw)),
               "pbf_side")
        return bool(uv.value), nw.value

    def rank(self, g, rank_type="ADEG"):
        """Returns (order, rank_of, W) as uint32 arrays (DeviceArrays for device graphs)."""
        order = self._alloc_out(g, g.n, np.uint32)
        rank_of = self._alloc_out(g, g.n, np.uint32)
        w = ctypes.c_int64()
        _check(_lib.pbf_rank(self._ctx, ctypes.byref(g._csr), self._mem(g), _enum(RANK_TYPES, rank_type, "rank type"),
                             self._ptr(order) if g.n else None, self._ptr(rank_of) if g.n else None, ctypes.byref(w)),
               "pbf_rank")
        return order, rank_of, w.value

    def count(self, g, use_v, per="VERT", count_type="GPU", rank_type="ADEG", out=None):
        """Count / CountE / CountTotal.  VERT -> uint64[n_S]; EDGE -> uint64[m]; TOTAL -> int."""
        per = _enum(PER_TYPES, per, "per type")
        ct = _enum(COUNT_TYPES, count_type, "count type")
        rt = _enum(RANK_TYPES, rank_type, "rank type")
        mem = self._mem(g)
        if per == 2:
            tot = np.zeros(1, dtype=np.uint64)
            if mem == MEM_DEVICE:
                dt = DeviceArray(self, 1, np.uint64)
                _check(_lib.pbf_count_total(self._ctx, ctypes.byref(g._csr), mem, int(bool(use_v)), ct, rt,
                                            dt.ptr), "pbf_count_total")
                return int(dt.numpy()[0])
            _check(_lib.pbf_count_total(self._ctx, ctypes.byref(g._csr), mem, int(bool(use_v)), ct, rt,
                                        tot.ctypes.data), "pbf_count_total")
            return int(tot[0])
        n_out = (g.nu if use_v else g.nv) if per == 0 else g.m
        if out is None:
            out = self._alloc_out(g, n_out, np.uint64)
        fn = _lib.pbf_count_vert if per == 0 else _lib.pbf_count_edge
        _check(fn(self._ctx, ctypes.byref(g._csr), mem, int(bool(use_v)), ct, rt, self._ptr(out) if n_out else 1),
               "pbf_count_vert" if per == 0 else "pbf_count_edge")
        return out

    def count_vert_noalloc(self, g, use_v, count_type="GPU", rank_type="ADEG", out=None):
        """Count per vertex through pbf_count_vert_noalloc (device graph and output; the
        library's fixed per-device scratch, no allocation inside the call)."""
        if not isinstance(g, DeviceGraph):
            raise TypeError("pbf_count_vert_noalloc takes device graphs")
        n_out = g.nu if use_v else g.nv
        if out is None:
            out = DeviceArray(self, n_out, np.uint64)
        _check(_lib.pbf_count_vert_noalloc(self._ctx, ctypes.byref(g._csr), int(bool(use_v)),
                                           _enum(COUNT_TYPES, count_type, "count type"),
                                           _enum(RANK_TYPES, rank_type, "rank type"), out.ptr if n_out else 1),
               "pbf_count_vert_noalloc")
        return out

    def edge_to_idxs(self, g, use_v):
        eti = self._alloc_out(g, g.m, np.uint32)
        _check(_lib.pbf_edge_to_idxs(self._ctx, ctypes.byref(g._csr), self._mem(g), int(bool(use_v)),
                                     self._ptr(eti) if g.m else 1), "pbf_edge_to_idxs")
        return eti

    def peel(self, g, use_v, counts, peel_type="GPU"):
        """Peel: tip numbers on S from per-vertex counts."""
        n_s = g.nu if use_v else g.nv
        counts = self._as_u64(g, counts)
        tip = self._alloc_out(g, n_s, np.uint64)
        _check(_lib.pbf_peel_vert(self._ctx, ctypes.byref(g._csr), self._mem(g), int(bool(use_v)),
                                  _enum(PEEL_TYPES, peel_type, "peel type"), self._ptr(counts) if n_s else 1,
                                  self._ptr(tip) if n_s else 1), "pbf_peel_vert")
        return tip

    def peel_edge(self, g, use_v, counts, peel_type="GPU"):
        """PeelE: wing numbers from per-edge counts (O-CSR order)."""
        counts = self._as_u64(g, counts)
        wing = self._alloc_out(g, g.m, np.uint64)
        _check(_lib.pbf_peel_edge(self._ctx, ctypes.byref(g._csr), self._mem(g), int(bool(use_v)),
                                  _enum(PEEL_TYPES, peel_type, "peel type"), self._ptr(counts) if g.m else 1,
                                  self._ptr(wing) if g.m else 1), "pbf_peel_edge")
        return wing

    # ---- approximate totals by sparsification ----
    def sparsify(self, g, sparse_type="EDGE", denom=25, seed=0):
        """eSparseBipartite / clrSparseBipartite + delZeroDeg: a Graph for a host graph, a
        DeviceGraph for a device graph."""
        st = _enum(SPARSE_TYPES, sparse_type, "sparse type")
        bufs = [self._alloc_out(g, n, np.uint32) for n in (g.nv + 1, max(g.m, 1), g.nu + 1, max(g.m, 1))]
        nv, nu, m = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(_lib.pbf_sparsify(self._ctx, ctypes.byref(g._csr), self._mem(g), st, int(denom), int(seed),
                                 ctypes.byref(nv), ctypes.byref(nu), ctypes.byref(m), *[self._ptr(b) for b in bufs]),
               "pbf_sparsify")
        nv, nu, m = nv.value, nu.value, m.value
        if isinstance(g, DeviceGraph):
            return DeviceGraph._from_arrays(self, nv, nu, m, *bufs)
        return Graph(nv, nu, bufs[0][:nv + 1], bufs[1][:m], bufs[2][:nu + 1], bufs[3][:m])

    def count_total_sparse(self, g, sparse_type="EDGE", denom=25, seed=0, count_type="GPU", rank_type="ADEG"):
        """CountTotal of the sparsified graph with the driver's scaling: (approx, sparse_total)."""
        approx, tot = ctypes.c_int64(), ctypes.c_uint64()
        _check(_lib.pbf_count_total_sparse(self._ctx, ctypes.byref(g._csr), self._mem(g),
                                           _enum(SPARSE_TYPES, sparse_type, "sparse type"), int(denom), int(seed),
                                           _enum(COUNT_TYPES, count_type, "count type"),
                                           _enum(RANK_TYPES, rank_type, "rank type"), ctypes.byref(approx),
                                           ctypes.byref(tot)), "pbf_count_total_sparse")
        return approx.value, tot.value

    # ---- batches (no counterpart in the reference) ----
    BATCH_UNION, BATCH_PLAIN, BATCH_CAPTURE = 0, 1, 16

    def _batch_args(self, graphs):
        mems = {self._mem(g) for g in graphs}
        if len(mems) > 1:
            raise ValueError("a batch is either all host graphs or all device graphs")
        mem = mems.pop() if mems else MEM_HOST
        arr = (_CSR * max(len(graphs), 1))()
        for k, g in enumerate(graphs):
            arr[k] = g._csr
        cap = sum(max(g.nv, g.nu) for g in graphs)
        return mem, arr, cap

    def count_batch(self, graphs, rank_type="ADEG", mode="union", capture=False, outputs=None):
        """Per-vertex counts and totals of many graphs.  mode: "union" (one pass over the
        disjoint union) or "plain" (graph by graph).  Host graphs give numpy results
        (list of per-graph uint64 arrays, uint64 totals); device graphs give
        (DeviceArray vert, DeviceArray offsets, DeviceArray totals), reusing `outputs` (a
        previous such triple) when given, which a captured batch needs for its replays."""
        mem, arr, cap = self._batch_args(graphs)
        flags = {"union": self.BATCH_UNION, "plain": self.BATCH_PLAIN}[mode] | (self.BATCH_CAPTURE if capture else 0)
        rt = _enum(RANK_TYPES, rank_type, "rank type")
        cnt = len(graphs)
        if mem == MEM_DEVICE:
            if outputs is None:
                outputs = (DeviceArray(self, max(cap, 1), np.uint64), DeviceArray(self, cnt + 1, np.uint64),
                           DeviceArray(self, max(cnt, 1), np.uint64))
            vout, voff, tot = outputs
            _check(_lib.pbf_count_batch(self._ctx, cnt, arr, mem, rt, flags, vout.ptr, vout.n, voff.ptr, tot.ptr,
                                        MEM_DEVICE), "pbf_count_batch")
            return outputs
        vout = np.zeros(max(cap, 1), np.uint64)
        voff = np.zeros(cnt + 1, np.uint64)
        tot = np.zeros(max(cnt, 1), np.uint64)
        _check(_lib.pbf_count_batch(self._ctx, cnt, arr, mem, rt, flags, vout.ctypes.data, len(vout), voff.ctypes.data,
                                    tot.ctypes.data, MEM_HOST), "pbf_count_batch")
        return [vout[int(voff[k]):int(voff[k + 1])] for k in range(cnt)], tot[:cnt]

    def batch_autotune(self, graphs, rank_type="ADEG"):
        """Runs the batch autotuner on these graphs; returns (best candidate index, ms)."""
        mem, arr, cap = self._batch_args(graphs)
        cnt = len(graphs)
        if mem == MEM_DEVICE:
            vout, voff, tot = (DeviceArray(self, max(cap, 1), np.uint64), DeviceArray(self, cnt + 1, np.uint64),
                               DeviceArray(self, max(cnt, 1), np.uint64))
            ptrs = (vout.ptr, voff.ptr, tot.ptr)
        else:
            vout, voff, tot = np.zeros(max(cap, 1), np.uint64), np.zeros(cnt + 1, np.uint64), np.zeros(max(cnt, 1), np.uint64)
            ptrs = (vout.ctypes.data, voff.ctypes.data, tot.ctypes.data)
        best, ms = ctypes.c_int(), ctypes.c_float()
        _check(_lib.pbf_batch_autotune(self._ctx, cnt, arr, mem, _enum(RANK_TYPES, rank_type, "rank type"), ptrs[0],
                                       max(cap, 1), ptrs[1], ptrs[2], mem, ctypes.byref(best), ctypes.byref(ms)),
               "pbf_batch_autotune")
        return best.value, ms.value

    def set_tuning(self, light_max=0, light_work=1 << 14, blocks_per_cu=4):
        _check(_lib.pbf_set_tuning(self._ctx, light_max, light_work, blocks_per_cu), "pbf_set_tuning")

    # ---- learned selector (no counterpart in the reference) ----
    def batch_features(self, graphs):
        """float32 [len(graphs), MLP_IN] feature rows (see pbf_batch_features)."""
        mem, arr, _ = self._batch_args(graphs)
        F = np.zeros((max(len(graphs), 1), MLP_IN), np.float32)
        _check(_lib.pbf_batch_features(self._ctx, len(graphs), arr, mem, F.ctypes.data), "pbf_batch_features")
        return F[:len(graphs)]

    def mlp_train(self, params, X, T, epochs, lr=1e-3, seed=1):
        """Adam on float32 params (updated in place); returns the last epoch's mean loss."""
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        assert params.dtype == np.float32 and params.size == MLP_PARAMS and params.flags.c_contiguous
        loss = ctypes.c_float()
        _check(_lib.pbf_mlp_train(self._ctx, params.ctypes.data, X.ctypes.data, T.ctypes.data, len(X), int(epochs),
                                  float(lr), int(seed), ctypes.byref(loss)), "pbf_mlp_train")
        return loss.value

    def mlp_grad(self, params, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = ctypes.c_float()
        _check(_lib.pbf_mlp_grad(self._ctx, np.ascontiguousarray(params, np.float32).ctypes.data, X.ctypes.data,
                                 T.ctypes.data, len(X), g.ctypes.data, ctypes.byref(loss)), "pbf_mlp_grad")
        return g, loss.value

    def mlp_predict(self, params, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((max(len(X), 1), MLP_OUT), np.float32)
        _check(_lib.pbf_mlp_predict(self._ctx, np.ascontiguousarray(params, np.float32).ctypes.data, X.ctypes.data,
                                    len(X), Y.ctypes.data), "pbf_mlp_predict")
        return Y[:len(X)]

    @staticmethod
    def _model(model):
        if model is None:
            return None
        if not isinstance(model, SelectorModel):
            model = SelectorModel.from_dict(model)
        return ctypes.byref(model)

    def select(self, graphs, model=None):
        """pbf_select: the chosen variant (0..8) of every graph; model None = the fixed variant."""
        mem, arr, _ = self._batch_args(graphs)
        v = np.zeros(max(len(graphs), 1), np.int32)
        _check(_lib.pbf_select(self._ctx, len(graphs), arr, mem, self._model(model), v.ctypes.data), "pbf_select")
        return v[:len(graphs)]

    def count_batch_learned(self, graphs, model=None, outputs=None):
        """pbf_count_batch_learned: (outputs as count_batch gives them, chosen variants)."""
        mem, arr, cap = self._batch_args(graphs)
        cnt = len(graphs)
        var = np.zeros(max(cnt, 1), np.int32)
        if mem == MEM_DEVICE:
            if outputs is None:
                outputs = (DeviceArray(self, max(cap, 1), np.uint64), DeviceArray(self, cnt + 1, np.uint64),
                           DeviceArray(self, max(cnt, 1), np.uint64))
            vout, voff, tot = outputs
            _check(_lib.pbf_count_batch_learned(self._ctx, cnt, arr, mem, self._model(model), vout.ptr, vout.n, voff.ptr,
                                                tot.ptr, MEM_DEVICE, var.ctypes.data), "pbf_count_batch_learned")
            return outputs, var[:cnt]
        vout = np.zeros(max(cap, 1), np.uint64)
        voff = np.zeros(cnt + 1, np.uint64)
        tot = np.zeros(max(cnt, 1), np.uint64)
        _check(_lib.pbf_count_batch_learned(self._ctx, cnt, arr, mem, self._model(model), vout.ctypes.data, len(vout),
                                            voff.ctypes.data, tot.ctypes.data, MEM_HOST, var.ctypes.data),
               "pbf_count_batch_learned")
        return ([vout[int(voff[k]):int(voff[k + 1])] for k in range(cnt)], tot[:cnt]), var[:cnt]

    def _as_u64(self, g, counts):
        if isinstance(g, DeviceGraph):
            if isinstance(counts, DeviceArray):
                return counts
            return DeviceArray(self, len(counts), np.uint64, counts)
        return np.ascontiguousarray(as_u64(counts), dtype=np.uint64)


MLP_IN, MLP_HIDDEN, MLP_OUT, MLP_PARAMS = 16, 64, 9, 6288
VARIANT_FIXED = 3  # PBF_VARIANT_FIXED: ADEG with the default knobs
RANK_NAMES = {v: k for k, v in RANK_TYPES.items()}


def mlp_init(seed):
    """He-uniform float32 parameters of the selector MLP (see pbf_mlp_init)."""
    p = np.zeros(MLP_PARAMS, np.float32)
    _check(_lib.pbf_mlp_init(int(seed), p.ctypes.data), "pbf_mlp_init")
    return p


def variant_tuning(v):
    """(rank type name, light_max, light_work, blocks_per_cu) of selector variant v (0..8)."""
    r, a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(_lib.pbf_variant_tuning(int(v), ctypes.byref(r), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
           "pbf_variant_tuning")
    return RANK_NAMES[r.value], a.value, b.value, c.value


def as_u64(x):
    """DeviceArray or host array -> numpy uint64."""
    if isinstance(x, DeviceArray):
        return x.numpy().view(np.uint64)
    return np.asarray(x).astype(np.uint64, copy=False)
