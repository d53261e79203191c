"""GPU parity at the benchmark configurations' full sizes.

The graphs are too large for the oracle to count whole, so the checks use
size-independent properties: the device orientation equals the oracle's entry
for entry, and the count over a strided sample of the tasks (task id %
stride == offset; kc_count_sampled) equals the oracle's count over the same
tasks, bit for bit.  The sampled tasks are spread over the whole task range,
so they include the heaviest tasks of the degeneracy core.
"""
import os
import time

import numpy as np
import pytest

import kcgpu_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)


def _log(*a):
    print(*a, flush=True)


def _oriented_pair(G, mode):
    """(device handle, oracle oriented CSR) after checking that they agree."""
    rp, col = G.csr()
    rp = rp.astype(np.uint64)
    Og = G.orient(mode)
    orp, oc, key = Og.export()
    if mode == "degen":
        pr, mc = O.peel_priority(rp, col)
        assert (key == pr).all() and Og.info()["max_core"] == mc
        erp, ec = O.orient_c(rp, col, pr)
    else:
        erp, ec = O.orient_c(rp, col)
    assert np.array_equal(orp, erp) and np.array_equal(oc, ec)
    return Og, erp, ec


def test_chung_lu_11m_sampled_vs_oracle():
    """Chung-Lu n=1,700,000 m=11,100,000 gamma=2.1 max expected degree 35,000
    seed 2, degeneracy order, edge pivoting at k=5..8: every 30th edge task.
    Plus a node-task sample and a sample at another offset."""
    t0 = time.perf_counter()
    G = K.Graph.generate(K.GEN_CHUNG_LU, seed=2, n=1_700_000, m=11_100_000, gamma=2.1, max_degree=35000.0)
    assert G.info()["m"] == 11_100_000
    Og, orp, oc = _oriented_pair(G, "degen")
    _log("chung-lu oriented and checked", f"{time.perf_counter() - t0:.1f}s", Og.info())
    for k, gran, stride, off in ((5, "edge", 30, 0), (6, "edge", 30, 0), (7, "edge", 30, 0), (8, "edge", 30, 0),
                                 (8, "edge", 97, 41), (7, "node", 30, 11)):
        t = time.perf_counter()
        exp = O.count(orp, oc, k, "pivot_tails", gran, stride=stride, offset=off, threads=NT)
        tc = time.perf_counter() - t
        r = Og.count_sampled(k, stride, off, "pivot", gran)
        _log(f"k={k} {gran} stride={stride} offset={off}: gpu={r.count} oracle={exp['count']} "
             f"tasks {r.raw['tasks']}/{exp['tasks']} gpu {r.raw['count_ms']:.1f} ms cpu {tc:.1f}s")
        assert r.count == exp["count"] and not r.overflow and not exp["overflow"], (k, gran, stride, off)
        assert r.raw["tasks"] == exp["tasks"], (k, gran, stride, off)
        assert r.raw["alg_bytes"] == exp["bytes"], (k, gran, stride, off)
    # the sampled shares add up to the whole count (three offsets of stride 3)
    full = Og.count(6, "pivot", "edge").count
    parts = [Og.count_sampled(6, 3, o, "pivot", "edge").count for o in range(3)]
    assert sum(parts) == full


def test_sampled_counts_partition_the_tasks(golden):
    """Samples at every offset of a stride add up to kc_count, for both
    granularities and algorithms, and each equals the oracle's sample."""
    g = next(x for x in golden if x["name"] == "planted_400")
    e = np.array(g["edges"], dtype=np.uint32).reshape(-1, 2)
    G = K.Graph.from_edges(e[:, 0], e[:, 1], g["n"])
    Og, orp, oc = _oriented_pair(G, "degen")
    for k in (1, 2, 3, 5, 7):
        for gran in ("node", "edge"):
            for algo in ("orient", "pivot"):
                parts = []
                for off in range(5):
                    r = Og.count_sampled(k, 5, off, algo, gran, part_size=4)
                    exp = O.count(orp, oc, k, "pivot", gran, stride=5, offset=off)
                    assert r.count == exp["count"], (k, gran, algo, off)
                    parts.append(r.count)
                assert sum(parts) == g["counts"][k], (k, gran, algo)
    with pytest.raises(K.KCError) as ei:
        Og.count_sampled(4, 5, 5)
    assert ei.value.code == K.KC_ERR_ARG
    with pytest.raises(K.KCError) as ei:
        Og.count_sampled(4, 0, 0)
    assert ei.value.code == K.KC_ERR_ARG
