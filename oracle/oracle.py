"""CPU oracle for the k-clique path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module; the product library never does.

Numpy restatements (reference file:line in each docstring):
  * :func:`csr_from_edges`  -- include/CSRCOO.cuh:37-80 (+ dedupe/symmetrize)
  * :func:`orient`          -- include/main_support.cuh:54-167, src/main.cu:179-209
C restatements (``kc_oracle.c``, loaded with ctypes):
  * :func:`peel_priority`   -- kcore/kcore.cuh:442-584
  * :func:`count`           -- kclique/kckernels_orientation.cuh, kckernels_pivoting.cuh,
                               kckernels.cuh:2263-2630
  * :func:`count_local`     -- kclique/kclique_local.cuh:24-639 (``-m lkc``)
  * :func:`binom`, :func:`brute_force`
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE])
        lib = ctypes.CDLL(path)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.orc_binom.restype = ctypes.c_uint64
        lib.orc_binom.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ip]
        lib.orc_peel_priority.restype = ctypes.c_int64
        lib.orc_peel_priority.argtypes = [ctypes.c_uint32, u64p, u32p, u32p]
        lib.orc_orient.restype = ctypes.c_uint64
        lib.orc_orient.argtypes = [ctypes.c_uint32, u64p, u32p, u32p, u64p, u32p]
        lib.orc_count.restype = ctypes.c_int
        lib.orc_count.argtypes = [ctypes.c_uint32, u64p, u32p, ctypes.c_uint32, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                  u64p, ip, u64p, u64p]
        lib.orc_count_local.restype = ctypes.c_int
        lib.orc_count_local.argtypes = [ctypes.c_uint32, u64p, u32p, ctypes.c_uint32, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, u64p, ip]
        lib.orc_brute_force.restype = ctypes.c_uint64
        lib.orc_brute_force.argtypes = [ctypes.c_uint32, u64p, ctypes.c_uint32, ip]
        lib.orc_threads.restype = ctypes.c_int
        lib.orc_threads.argtypes = []
        _LIB = lib
    return _LIB


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def csr_from_edges(src, dst, n=None):
    """Symmetric simple CSR from an edge list.

    Restates ``CSRCOO::from_edgelist`` (include/CSRCOO.cuh:37-80): self-loops
    dropped, rows sorted by source, columns sorted inside a row (the reference
    needs a src-sorted input, include/FIleReader.cuh:314-320).  Unlike the
    reference it symmetrizes and removes duplicate entries, so the result is
    the CSR of the simple undirected graph.  ``n`` defaults to max id + 1.
    Returns ``(rowptr uint64[n+1], col uint32[nnz])``.
    """
    src = np.asarray(src, dtype=np.uint64)
    dst = np.asarray(dst, dtype=np.uint64)
    keep = src != dst
    s = np.concatenate([src[keep], dst[keep]])
    d = np.concatenate([dst[keep], src[keep]])
    if n is None:
        n = int(max(s.max(initial=0), d.max(initial=0)) + 1) if s.size else 0
    key = np.unique((s << np.uint64(32)) | d)
    rows = (key >> np.uint64(32)).astype(np.int64)
    col = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    rowptr = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.uint64)
    return rowptr, col


def orient(rowptr, col, key=None):
    """Oriented CSR (src/main.cu:179-209).

    keep(u->v) <=> (key u, u) < (key v, v); key defaults to the full degree
    (include/main_support.cuh:54-86 with ``getNodeDegree``); a priority array
    gives the degeneracy rule (:89-128).  Kept entries stay in row order
    (stable compaction, CUBSelect) and the row pointer is the exclusive scan
    of the per-row kept count (warp_detect_deleted_edges, :133-167).
    Returns ``(rowptr uint64[n+1], col uint32[m])``.
    """
    rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    n = rowptr.size - 1
    deg = np.diff(rowptr).astype(np.int64)
    k = deg if key is None else np.asarray(key, dtype=np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    dst = col.astype(np.int64)
    keep = (k[src] < k[dst]) | ((k[src] == k[dst]) & (src < dst))
    cnt = np.bincount(src[keep], minlength=n)
    orp = np.zeros(n + 1, dtype=np.uint64)
    orp[1:] = np.cumsum(cnt)
    return orp, col[keep].copy()


def orient_c(rowptr, col, key=None):
    """Same as :func:`orient`, by the C restatement ``orc_orient``."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    n = rowptr.size - 1
    orp = np.zeros(n + 1, dtype=np.uint64)
    oc = np.zeros(max(col.size // 2, 1), dtype=np.uint32)
    kp = None
    if key is not None:
        key = np.ascontiguousarray(key, dtype=np.uint32)
        kp = _p(key, ctypes.c_uint32)
    m = _lib().orc_orient(n, _p(rowptr, ctypes.c_uint64), _p(col, ctypes.c_uint32), kp,
                          _p(orp, ctypes.c_uint64), _p(oc, ctypes.c_uint32))
    return orp, oc[:m].copy()


def peel_priority(rowptr, col):
    """Degeneracy key of kcore/kcore.cuh:442-584 (see kc_oracle.c).

    Returns ``(priority uint32[n], max_core)``."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    n = rowptr.size - 1
    prio = np.zeros(max(n, 1), dtype=np.uint32)
    mc = _lib().orc_peel_priority(n, _p(rowptr, ctypes.c_uint64), _p(col, ctypes.c_uint32),
                                  _p(prio, ctypes.c_uint32))
    return prio[:n], int(mc)


ALGO = {"orient": 0, "pivot": 1, "pivot_tails": 2}
GRAN = {"node": 0, "edge": 1}


def count(orowptr, ocol, k, algo="pivot", gran="node", stride=1, offset=0, threads=0):
    """k-clique count over an oriented CSR by one of the four traversals.

    Returns a dict with ``count``, ``overflow``, ``tasks`` and ``bytes``
    (algorithmic bytes of the processed tasks)."""
    orowptr = np.ascontiguousarray(orowptr, dtype=np.uint64)
    ocol = np.ascontiguousarray(ocol, dtype=np.uint32)
    n = orowptr.size - 1
    c = ctypes.c_uint64(0)
    o = ctypes.c_int(0)
    t = ctypes.c_uint64(0)
    b = ctypes.c_uint64(0)
    rc = _lib().orc_count(n, _p(orowptr, ctypes.c_uint64), _p(ocol, ctypes.c_uint32), int(k),
                          ALGO[algo], GRAN[gran], int(stride), int(offset), int(threads),
                          ctypes.byref(c), ctypes.byref(o), ctypes.byref(t), ctypes.byref(b))
    if rc != 0:
        raise ValueError("orc_count failed")
    return {"count": c.value, "overflow": bool(o.value), "tasks": t.value, "bytes": b.value}


def count_local(orowptr, ocol, k, algo="pivot", gran="node", threads=0):
    """Per-vertex k-clique counts (``-m lkc``): (uint64 array of n, overflow)."""
    orowptr = np.ascontiguousarray(orowptr, dtype=np.uint64)
    ocol = np.ascontiguousarray(ocol, dtype=np.uint32)
    n = orowptr.size - 1
    out = np.zeros(max(n, 1), dtype=np.uint64)
    o = ctypes.c_int(0)
    rc = _lib().orc_count_local(n, _p(orowptr, ctypes.c_uint64), _p(ocol, ctypes.c_uint32), int(k),
                                ALGO[algo], GRAN[gran], int(threads), _p(out, ctypes.c_uint64), ctypes.byref(o))
    if rc != 0:
        raise ValueError("orc_count_local failed")
    return out[:n], bool(o.value)


def binom(n, r):
    o = ctypes.c_int(0)
    v = _lib().orc_binom(int(n), int(r), ctypes.byref(o))
    return v, bool(o.value)


def brute_force(n, src, dst, k):
    """k-cliques of a graph with n <= 64 vertices by enumerating subsets."""
    if n > 64:
        raise ValueError("brute force needs n <= 64")
    adj = np.zeros(max(n, 1), dtype=np.uint64)
    for a, b in zip(src, dst):
        a = int(a); b = int(b)
        if a == b:
            continue
        adj[a] |= np.uint64(1) << np.uint64(b)
        adj[b] |= np.uint64(1) << np.uint64(a)
    o = ctypes.c_int(0)
    v = _lib().orc_brute_force(n, _p(adj, ctypes.c_uint64), int(k), ctypes.byref(o))
    return v, bool(o.value)


def threads():
    return int(_lib().orc_threads())
