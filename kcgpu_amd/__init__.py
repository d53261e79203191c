"""kcgpu_amd -- Python host mirror of the kcgpu C ABI (include/kc.h).

The compute path is the native library ``libkcgpu.so`` (HIP kernels for
gfx950 behind a C ABI); this module only binds it with ctypes.  There is no
Python or CPU fallback: importing fails loudly when the library is missing.

The names follow the reference's drop-in surface:
``Graph.load`` (EdgeListFile + CSRCOO), ``Graph.orient`` (-o degree|degen),
``Oriented.count`` (-m kc with -k, -p node|edge, -q [o|p]<P>[b|n]).
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkcgpu.so")

KC_OK = 0
KC_ERR_ARG = 1
KC_ERR_IO = 2
KC_ERR_FORMAT = 3
KC_ERR_DEVICE = 4
KC_ERR_NOMEM = 5
KC_ERR_OVERFLOW = 6
KC_ERR_BUDGET = 7
KC_ERR_NOTFOUND = 8
KC_ERR_LIMIT = 9

NOALLOC_SCRATCH_MIB = 256  # KC_NOALLOC_SCRATCH_MIB: size of kc_count_noalloc's fixed device array

GEN_ER, GEN_CHUNG_LU, GEN_RMAT, GEN_BA = 0, 1, 2, 3
ORIENT = {"degree": 0, "degen": 1}
ALGO = {"orient": 0, "pivot": 1}
GRAN = {"node": 0, "edge": 1}


class KCError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kcgpu error {code}: {msg}")
        self.code = code


class GraphInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("nnz", ctypes.c_uint64), ("max_degree", ctypes.c_uint32),
                ("checksum", ctypes.c_uint64)]


class GenParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_uint32), ("m", ctypes.c_uint64),
                ("gamma", ctypes.c_double), ("max_degree", ctypes.c_double), ("scale", ctypes.c_uint32),
                ("edge_factor", ctypes.c_uint32), ("a", ctypes.c_double), ("b", ctypes.c_double),
                ("c", ctypes.c_double), ("attach", ctypes.c_uint32), ("seed", ctypes.c_uint64)]


class OrientedInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint64), ("max_out_degree", ctypes.c_uint32),
                ("max_core", ctypes.c_uint32), ("preprocess_ms", ctypes.c_double)]


class Query(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("algo", ctypes.c_uint32), ("gran", ctypes.c_uint32),
                ("part_size", ctypes.c_uint32), ("binary", ctypes.c_uint32), ("auto_tune", ctypes.c_uint32)]


class Result(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("overflow", ctypes.c_uint32), ("algo", ctypes.c_uint32),
                ("gran", ctypes.c_uint32), ("part_size", ctypes.c_uint32), ("tasks", ctypes.c_uint64),
                ("alg_bytes", ctypes.c_uint64), ("tree_nodes", ctypes.c_uint64), ("count_ms", ctypes.c_double),
                ("dominant_ms", ctypes.c_double), ("dominant_bytes", ctypes.c_uint64),
                ("dominant_class", ctypes.c_uint32), ("class_ms", ctypes.c_double * 6),
                ("class_bytes", ctypes.c_uint64 * 6), ("class_tasks", ctypes.c_uint64 * 6),
                ("class_kernel", ctypes.c_uint32 * 6), ("class_max_task_ms", ctypes.c_double * 6),
                ("class_task_ms_sum", ctypes.c_double * 6)]

    def as_dict(self):
        d = {}
        for f in self._fields_:
            v = getattr(self, f[0])
            d[f[0]] = list(v) if hasattr(v, "_length_") else v
        return d


class Plan(ctypes.Structure):
    _fields_ = [("orient_mode", ctypes.c_int), ("algo", ctypes.c_uint32), ("gran", ctypes.c_uint32),
                ("part_size", ctypes.c_uint32), ("degree_max_out", ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_uint64), ("workers", ctypes.c_uint32), ("max_batch", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("mean_batch", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("queries_per_s", ctypes.c_double), ("graphs", ctypes.c_uint64),
                ("resident_bytes", ctypes.c_uint64), ("held_scratch_bytes", ctypes.c_uint64),
                ("resident_admitted", ctypes.c_uint64), ("joined_running", ctypes.c_uint64),
                ("resident_launches", ctypes.c_uint64)]


class BatchStats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("count_ms", ctypes.c_double), ("tasks", ctypes.c_uint64),
                ("queries", ctypes.c_uint64)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32), ("cu_count", ctypes.c_int),
                ("hbm_bytes", ctypes.c_uint64), ("lds_per_block", ctypes.c_uint64), ("clock_khz", ctypes.c_int)]


# every symbol include/kc.h declares, with (restype, argtypes)
_vp = ctypes.c_void_p
_u32p = ctypes.POINTER(ctypes.c_uint32)
SIGNATURES = {
    "kc_last_error": (ctypes.c_char_p, []),
    "kc_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint, ctypes.POINTER(_vp)]),
    "kc_read_edges_reference": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_u32p), ctypes.POINTER(_u32p),
                                               ctypes.POINTER(ctypes.c_size_t)]),
    "kc_free_buffer": (None, [_vp]),
    "kc_graph_from_edges": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_size_t, _u32p, _u32p, ctypes.POINTER(_vp)]),
    "kc_graph_from_csr": (ctypes.c_int, [ctypes.c_uint32, _u32p, _u32p, ctypes.POINTER(_vp)]),
    "kc_graph_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(GraphInfo)]),
    "kc_graph_get_csr": (ctypes.c_int, [_vp, _u32p, _u32p]),
    "kc_graph_free": (None, [_vp]),
    "kc_generate": (ctypes.c_int, [ctypes.POINTER(GenParams), ctypes.POINTER(_vp)]),
    "kc_orient": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "kc_orient_batch": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int), ctypes.c_size_t,
                                       ctypes.POINTER(_vp)]),
    "kc_orient_edges": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_size_t, _u32p, _u32p, ctypes.c_int,
                                       ctypes.POINTER(_vp)]),
    "kc_graph_from_edges_device": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_size_t, _u32p, _u32p,
                                                  ctypes.POINTER(_vp)]),
    "kc_oriented_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(OrientedInfo)]),
    "kc_oriented_export": (ctypes.c_int, [_vp, _u32p, _u32p, _u32p]),
    "kc_oriented_free": (None, [_vp]),
    "kc_oriented_save": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "kc_oriented_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "kc_convert": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint]),
    "kc_count": (ctypes.c_int, [_vp, ctypes.POINTER(Query), ctypes.POINTER(Result)]),
    "kc_count_sampled": (ctypes.c_int, [_vp, ctypes.POINTER(Query), ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.POINTER(Result)]),
    "kc_count_local": (ctypes.c_int, [_vp, ctypes.POINTER(Query), ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(Result)]),
    "kc_count_noalloc": (ctypes.c_int, [_vp, ctypes.POINTER(Query), ctypes.POINTER(Result)]),
    "kc_parse_spec": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(Query)]),
    "kc_get_device_info": (ctypes.c_int, [ctypes.POINTER(DeviceInfo)]),
    "kc_select_device": (ctypes.c_int, [ctypes.c_int]),
    "kc_shutdown": (ctypes.c_int, []),
    "kc_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]),
    "kc_server_add_graph": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _u32p]),
    "kc_server_remove_graph": (ctypes.c_int, [_vp, ctypes.c_uint32]),
    "kc_server_submit": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(Query), ctypes.POINTER(ctypes.c_uint64)]),
    "kc_server_poll": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(Result), ctypes.POINTER(ctypes.c_int)]),
    "kc_server_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(Result)]),
    "kc_server_drain": (ctypes.c_int, [_vp]),
    "kc_server_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ServerStats)]),
    "kc_server_reset_stats": (ctypes.c_int, [_vp]),
    "kc_server_destroy": (None, [_vp]),
    "kc_autotune": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(Plan)]),
    "kc_batch_count": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(Query), ctypes.c_size_t,
                                      ctypes.POINTER(Result), ctypes.POINTER(BatchStats)]),
}

_lib = None


def lib():
    """The loaded native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m kcgpu_amd.build` "
                              "(there is no CPU fallback)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


# Handles still alive at interpreter exit are freed, and the library's
# streams destroyed, from an atexit handler: that runs before the interpreter
# tears its objects down, so no HIP call is left for the HIP runtime's (or a
# profiler's) own exit handlers.
_live = weakref.WeakSet()
_shut = False


def shutdown():
    """Free every live handle and release the library's device state (kc_shutdown)."""
    global _shut
    if _lib is None or _shut:
        return
    for obj in list(_live):
        obj.close()
    _lib.kc_shutdown()
    _shut = True


atexit.register(shutdown)


def _check(rc):
    if rc != KC_OK:
        raise KCError(rc, lib().kc_last_error().decode())


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_u32p)


class Graph:
    """Simple undirected graph (host CSR) -- the reference's loaded COOCSRGraph."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        _live.add(self)

    def close(self):
        if getattr(self, "_h", None) and _lib is not None and not _shut:
            _lib.kc_graph_free(self._h)
        self._h = None

    __del__ = close

    @classmethod
    def load(cls, path):
        h = ctypes.c_void_p()
        _check(lib().kc_load(os.fsencode(path), 0, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, src, dst, n=0):
        s, sp = _u32(src)
        d, dp = _u32(dst)
        if s.size != d.size:
            raise ValueError("src and dst differ in length")
        h = ctypes.c_void_p()
        _check(lib().kc_graph_from_edges(int(n), s.size, sp, dp, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges_device(cls, src, dst, n=0):
        """from_edges with the CSR built on the GPU (kc_graph_from_edges_device)."""
        s, sp = _u32(src)
        d, dp = _u32(dst)
        if s.size != d.size:
            raise ValueError("src and dst differ in length")
        h = ctypes.c_void_p()
        _check(lib().kc_graph_from_edges_device(int(n), s.size, sp, dp, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csr(cls, rowptr, col):
        r, rp = _u32(rowptr)
        c, cp = _u32(col)
        h = ctypes.c_void_p()
        _check(lib().kc_graph_from_csr(r.size - 1, rp, cp, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def generate(cls, kind, seed, n=0, m=0, gamma=0.0, max_degree=0.0, scale=0, edge_factor=0,
                 a=0.0, b=0.0, c=0.0, attach=0):
        p = GenParams(kind, n, m, gamma, max_degree, scale, edge_factor, a, b, c, attach, seed)
        h = ctypes.c_void_p()
        _check(lib().kc_generate(ctypes.byref(p), ctypes.byref(h)))
        return cls(h.value)

    def info(self):
        i = GraphInfo()
        _check(lib().kc_graph_get_info(self._h, ctypes.byref(i)))
        return {"n": i.n, "nnz": i.nnz, "m": i.nnz // 2, "max_degree": i.max_degree, "checksum": i.checksum}

    def csr(self):
        inf = self.info()
        rp = np.zeros(inf["n"] + 1, dtype=np.uint32)
        col = np.zeros(max(inf["nnz"], 1), dtype=np.uint32)
        _check(lib().kc_graph_get_csr(self._h, rp.ctypes.data_as(_u32p), col.ctypes.data_as(_u32p)))
        return rp, col[: inf["nnz"]]

    def autotune(self, k):
        """The library's plan: dict(orient, algo, gran, part_size, degree_max_out)."""
        p = Plan()
        _check(lib().kc_autotune(self._h, int(k), ctypes.byref(p)))
        return {"orient": "degen" if p.orient_mode == 1 else "degree", "algo": "pivot" if p.algo else "orient",
                "gran": "edge" if p.gran else "node", "part_size": p.part_size, "degree_max_out": p.degree_max_out}

    def orient(self, mode="degree"):
        h = ctypes.c_void_p()
        _check(lib().kc_orient(self._h, ORIENT[mode] if isinstance(mode, str) else int(mode), ctypes.byref(h)))
        return Oriented(h.value)


def orient_edges(src, dst, mode="degree", n=0):
    """Oriented graph of an edge list, with the CSR built and oriented on the
    GPU (kc_orient_edges): no host CSR."""
    s, sp = _u32(src)
    d, dp = _u32(dst)
    if s.size != d.size:
        raise ValueError("src and dst differ in length")
    h = ctypes.c_void_p()
    _check(lib().kc_orient_edges(int(n), s.size, sp, dp, ORIENT[mode], ctypes.byref(h)))
    return Oriented(h.value)


def orient_batch(graphs, modes):
    """Orient many graphs at once (kc_orient_batch): list of Oriented."""
    n = len(graphs)
    gs = (_vp * max(n, 1))(*[g._h.value for g in graphs])
    ms = (ctypes.c_int * max(n, 1))(*[ORIENT[m] if isinstance(m, str) else int(m) for m in modes])
    out = (_vp * max(n, 1))()
    _check(lib().kc_orient_batch(gs, ms, n, out))
    return [Oriented(out[i]) for i in range(n)]


def read_edges_reference(path):
    """The reference's raw read loop (defects included): (src, dst) arrays."""
    s = _u32p()
    d = _u32p()
    n = ctypes.c_size_t()
    _check(lib().kc_read_edges_reference(os.fsencode(path), ctypes.byref(s), ctypes.byref(d), ctypes.byref(n)))
    try:
        src = np.ctypeslib.as_array(s, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint32)
        dst = np.ctypeslib.as_array(d, shape=(n.value,)).copy() if n.value else np.zeros(0, np.uint32)
    finally:
        lib().kc_free_buffer(ctypes.cast(s, ctypes.c_void_p))
        lib().kc_free_buffer(ctypes.cast(d, ctypes.c_void_p))
    return src, dst


CONVERSIONS = {"mtx-bel": 0, "tsv-bel": 1, "tsv-mtx": 2, "bel-mtx": 3, "txt-bel": 4}
KC_CONV_REFERENCE_FLAGS = 1


def convert(kind, src, dst, reference_flags=False):
    """-m mtx-bel | tsv-bel | tsv-mtx | bel-mtx | txt-bel: write dst from src."""
    _check(lib().kc_convert(CONVERSIONS[kind] if isinstance(kind, str) else int(kind), os.fsencode(src),
                            os.fsencode(dst), KC_CONV_REFERENCE_FLAGS if reference_flags else 0))


def parse_spec(spec):
    """-q grammar [o|p]<P>[b|n|h] -> dict(algo, part_size, binary)."""
    q = Query()
    _check(lib().kc_parse_spec(spec.encode(), ctypes.byref(q)))
    return {"algo": "pivot" if q.algo == 1 else "orient", "part_size": q.part_size, "binary": bool(q.binary)}


@dataclass
class CountResult:
    count: int
    overflow: bool
    raw: dict


class Oriented:
    """Oriented CSR resident in HBM."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        _live.add(self)

    def close(self):
        if getattr(self, "_h", None) and _lib is not None and not _shut:
            _lib.kc_oriented_free(self._h)
        self._h = None

    __del__ = close

    def info(self):
        i = OrientedInfo()
        _check(lib().kc_oriented_get_info(self._h, ctypes.byref(i)))
        return {"n": i.n, "m": i.m, "max_out_degree": i.max_out_degree, "max_core": i.max_core,
                "preprocess_ms": i.preprocess_ms}

    def export(self):
        inf = self.info()
        rp = np.zeros(inf["n"] + 1, dtype=np.uint32)
        col = np.zeros(max(inf["m"], 1), dtype=np.uint32)
        key = np.zeros(max(inf["n"], 1), dtype=np.uint32)
        _check(lib().kc_oriented_export(self._h, rp.ctypes.data_as(_u32p), col.ctypes.data_as(_u32p),
                                        key.ctypes.data_as(_u32p)))
        return rp, col[: inf["m"]], key[: inf["n"]]

    def save(self, path):
        """Write the binary oriented-CSR cache file."""
        _check(lib().kc_oriented_save(self._h, os.fsencode(path)))

    @staticmethod
    def load(path):
        """Load a cache file written by save() straight into HBM."""
        h = ctypes.c_void_p()
        _check(lib().kc_oriented_load(os.fsencode(path), ctypes.byref(h)))
        return Oriented(h.value)

    def count(self, k, algo="pivot", gran="node", part_size=8, binary=True, auto_tune=False, check=True):
        q = Query(int(k), ALGO[algo] if isinstance(algo, str) else int(algo),
                  GRAN[gran] if isinstance(gran, str) else int(gran), int(part_size), int(bool(binary)),
                  int(bool(auto_tune)))
        r = Result()
        rc = lib().kc_count(self._h, ctypes.byref(q), ctypes.byref(r))
        if rc not in (KC_OK, KC_ERR_OVERFLOW) or (check and rc != KC_OK):
            _check(rc)
        return CountResult(int(r.count), bool(r.overflow), r.as_dict())

    def count_sampled(self, k, stride, offset=0, algo="pivot", gran="node", part_size=1, binary=True,
                      auto_tune=False, check=True):
        """kc_count over the tasks whose id % stride == offset (vertex ids, or
        oriented edges in CSR order): the sample a CPU enumerator can also count."""
        q = _query(k, algo, gran, part_size, binary, auto_tune)
        r = Result()
        rc = lib().kc_count_sampled(self._h, ctypes.byref(q), int(stride), int(offset), ctypes.byref(r))
        if rc not in (KC_OK, KC_ERR_OVERFLOW) or (check and rc != KC_OK):
            _check(rc)
        return CountResult(int(r.count), bool(r.overflow), r.as_dict())


def _count_noalloc(self, k, algo="pivot", gran="node", part_size=1, binary=True, auto_tune=False, check=True):
    """kc_count_noalloc: the same count with no device allocation (the launch
    scratch is the library's fixed device array; timing fields stay 0)."""
    q = _query(k, algo, gran, part_size, binary, auto_tune)
    r = Result()
    rc = lib().kc_count_noalloc(self._h, ctypes.byref(q), ctypes.byref(r))
    if rc not in (KC_OK, KC_ERR_OVERFLOW) or (check and rc != KC_OK):
        _check(rc)
    return CountResult(int(r.count), bool(r.overflow), r.as_dict())


Oriented.count_noalloc = _count_noalloc


def _count_local(self, k, algo="pivot", gran="node", part_size=8, binary=True, auto_tune=False, check=True):
    """-m lkc: (per-vertex uint64 counts, CountResult of the total)."""
    q = _query(k, algo, gran, part_size, binary, auto_tune)
    n = self.info()["n"]
    loc = np.zeros(max(n, 1), dtype=np.uint64)
    r = Result()
    rc = lib().kc_count_local(self._h, ctypes.byref(q), loc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                              ctypes.byref(r))
    if rc not in (KC_OK, KC_ERR_OVERFLOW) or (check and rc != KC_OK):
        _check(rc)
    return loc[:n], CountResult(int(r.count), bool(r.overflow), r.as_dict())


Oriented.count_local = _count_local


def _query(k, algo="pivot", gran="node", part_size=1, binary=True, auto_tune=False):
    return Query(int(k), ALGO[algo] if isinstance(algo, str) else int(algo),
                 GRAN[gran] if isinstance(gran, str) else int(gran), int(part_size), int(bool(binary)),
                 int(bool(auto_tune)))


def batch_count(pairs, check=True):
    """Count many (Oriented, query-dict) pairs in one persistent launch.

    ``pairs`` is a list of ``(oriented, dict(k=..., algo=..., gran=..., part_size=..., auto_tune=...))``.
    Returns ``(list of CountResult, stats dict)``."""
    n = len(pairs)
    hs = (_vp * max(n, 1))(*[o._h.value for o, _ in pairs])
    qs = (Query * max(n, 1))(*[_query(**q) for _, q in pairs])
    rs = (Result * max(n, 1))()
    st = BatchStats()
    rc = lib().kc_batch_count(hs, qs, n, rs, ctypes.byref(st))
    if rc not in (KC_OK, KC_ERR_OVERFLOW) or (check and rc != KC_OK):
        _check(rc)
    out = [CountResult(int(rs[i].count), bool(rs[i].overflow), rs[i].as_dict()) for i in range(n)]
    return out, {"total_ms": st.total_ms, "count_ms": st.count_ms, "tasks": st.tasks, "queries": st.queries}


class Server:
    """Serving loop: resident oriented graphs, queued (graph, k) queries,
    continuous batching on the GPU (kc_server_* in include/kc.h)."""

    def __init__(self, budget_bytes=0, workers=0, max_batch=0):
        self._h = ctypes.c_void_p()
        cfg = ServerConfig(int(budget_bytes), int(workers), int(max_batch))
        _check(lib().kc_server_create(ctypes.byref(cfg), ctypes.byref(self._h)))
        _live.add(self)

    def close(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None and not _shut:
            _lib.kc_server_destroy(self._h)
        self._h = ctypes.c_void_p()

    __del__ = close

    def add_graph(self, graph, orient="degen"):
        gid = ctypes.c_uint32()
        _check(lib().kc_server_add_graph(self._h, graph._h, ORIENT[orient], ctypes.byref(gid)))
        return gid.value

    def remove_graph(self, gid):
        _check(lib().kc_server_remove_graph(self._h, int(gid)))

    def submit(self, gid, k, algo="pivot", gran="node", part_size=1, auto_tune=False):
        q = _query(k, algo, gran, part_size, True, auto_tune)
        t = ctypes.c_uint64()
        _check(lib().kc_server_submit(self._h, int(gid), ctypes.byref(q), ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        r = Result()
        done = ctypes.c_int()
        rc = lib().kc_server_poll(self._h, int(ticket), ctypes.byref(r), ctypes.byref(done))
        if rc not in (KC_OK, KC_ERR_OVERFLOW):
            _check(rc)
        return CountResult(int(r.count), bool(r.overflow), r.as_dict()) if done.value else None

    def wait(self, ticket):
        r = Result()
        rc = lib().kc_server_wait(self._h, int(ticket), ctypes.byref(r))
        if rc not in (KC_OK, KC_ERR_OVERFLOW):
            _check(rc)
        return CountResult(int(r.count), bool(r.overflow), r.as_dict())

    def drain(self):
        _check(lib().kc_server_drain(self._h))

    def stats(self):
        st = ServerStats()
        _check(lib().kc_server_get_stats(self._h, ctypes.byref(st)))
        return {f[0]: getattr(st, f[0]) for f in st._fields_}

    def reset_stats(self):
        _check(lib().kc_server_reset_stats(self._h))


def device_info():
    d = DeviceInfo()
    _check(lib().kc_get_device_info(ctypes.byref(d)))
    return {"name": d.name.decode(), "arch": d.arch.decode(), "cu_count": d.cu_count, "hbm_bytes": d.hbm_bytes,
            "lds_per_block": d.lds_per_block, "clock_khz": d.clock_khz}
