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


def test_batched_mix_4096_graphs_full_size():
    """The batched workload at full size: 4,096 graphs (even i Barabasi-Albert
    n~U[200,20000] attach~U[3,20], odd i Erdos-Renyi n~U[200,20000] average
    degree~U[4,40], seed(i) = splitmix64(4^i), k~U{4..7}), autotuned plans,
    batched orientation.  One kc_batch_count over all of them equals the
    per-call kc_count of every graph, every 64th graph equals the oracle, and
    the batch is at least 3x the per-call throughput."""
    from concurrent.futures import ThreadPoolExecutor
    from kcgpu_amd.workloads import batch_mix, make_graph

    specs = batch_mix(4096)

    def make(sp):
        G = make_graph(K, sp)
        return G, G.autotune(sp["k"])

    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=NT) as ex:
        made = list(ex.map(make, specs))
    ors = K.orient_batch([G for G, _ in made], [p["orient"] for _, p in made])
    _log(f"4096 graphs generated and oriented in {time.perf_counter() - t:.1f}s, "
         f"{sum(o.info()['m'] for o in ors)} oriented edges")
    queries = [dict(k=sp["k"], algo=p["algo"], gran=p["gran"], part_size=p["part_size"])
               for sp, (_, p) in zip(specs, made)]
    ors[0].count(**queries[0])
    K.batch_count([(ors[0], queries[0])])
    t = time.perf_counter()
    plain = [o.count(**q).count for o, q in zip(ors, queries)]
    t_plain = time.perf_counter() - t
    t = time.perf_counter()
    res, st = K.batch_count(list(zip(ors, queries)))
    t_batch = time.perf_counter() - t
    _log(f"per-call {t_plain:.3f}s, one batch {t_batch:.3f}s ({t_plain / t_batch:.1f}x), {st}")
    assert [r.count for r in res] == plain
    assert t_plain / t_batch >= 3.0
    for i in range(0, 4096, 64):
        rp, col, _ = ors[i].export()
        q = queries[i]
        exp = O.count(rp.astype(np.uint64), col, q["k"], "pivot_tails", q["gran"], threads=NT)["count"]
        assert plain[i] == exp, (i, specs[i], q)
    # mixed algorithms and partition sizes inside one batch: each query keeps its own
    mix = []
    for i in range(256):
        q = dict(queries[i])
        if i % 3 == 1:
            q.update(algo="orient", part_size=(1, 8, 64)[i % 9 // 3])
        elif i % 3 == 2:
            q.update(gran="edge")
        mix.append((ors[i], q))
    res2, _ = K.batch_count(mix)
    for (o, q), r in zip(mix, res2):
        assert r.count == o.count(**q).count, q


RMAT = dict(edge_factor=16, a=0.57, b=0.19, c=0.19)


def _rmat22_samples(cases, full_ks=()):
    G = K.Graph.generate(K.GEN_RMAT, seed=3, scale=22, **RMAT)
    _log("rmat s22", G.info())
    full = {}
    for mode, cs in cases.items():
        Og, orp, oc = _oriented_pair(G, mode)
        _log(mode, Og.info())
        for k, gran, stride, off in cs:
            t = time.perf_counter()
            exp = O.count(orp, oc, k, "pivot_tails", gran, stride=stride, offset=off, threads=NT)
            tc = time.perf_counter() - t
            r = Og.count_sampled(k, stride, off, "pivot", gran)
            _log(f"{mode} k={k} {gran} stride={stride} offset={off}: gpu={r.count} oracle={exp['count']} "
                 f"tasks {r.raw['tasks']} gpu {r.raw['count_ms']:.1f} ms cpu {tc:.1f}s")
            assert r.count == exp["count"] and not r.overflow and not exp["overflow"], (mode, k, gran, stride)
            assert r.raw["tasks"] == exp["tasks"]
        for k in full_ks:
            for gran in ("node", "edge"):
                full[(mode, gran, k)] = Og.count(k, "pivot", gran).count
        del Og
    return full


def test_rmat_scale22_sampled_vs_oracle():
    """R-MAT scale 22, edge factor 16, (a, b, c, d) = (.57, .19, .19, .05),
    seed 3, deduplicated, symmetrized, loops dropped (64.2 M edges, degeneracy
    961), degeneracy order: the orientation equals the oracle's entry for
    entry, and strided task samples at k = 3..7, node and edge tasks, equal
    the oracle's counts over the same tasks."""
    _rmat22_samples({"degen": ((3, "edge", 97, 0), (4, "edge", 97, 1), (5, "node", 97, 2), (6, "node", 499, 3),
                               (6, "edge", 4999, 4), (7, "node", 2003, 5), (7, "edge", 49999, 6))})


def test_rmat_scale22_k8_k9_degree_order_and_modes():
    """The same R-MAT scale 22 graph: edge-task samples at k = 8 and 9
    (degeneracy order), samples in the degree order, and the full counts at
    k = 3..5 agree across the four (orientation, granularity) modes."""
    full = _rmat22_samples({"degen": ((8, "edge", 199999, 7), (9, "edge", 999983, 9)),
                            "degree": ((4, "node", 97, 7), (5, "edge", 997, 8), (6, "node", 997, 9))},
                           full_ks=(3, 4, 5))
    for k in (3, 4, 5):
        vals = {v for (_, _, kk), v in full.items() if kk == k}
        _log(f"k={k} full counts over four modes: {vals}")
        assert len(vals) == 1, (k, full)


def test_rmat_scale16_k3_to_11_vs_oracle():
    """R-MAT scale 16 (same generator and parameters as the scale-22 graph,
    seed 3; 909,144 edges, degeneracy 217), degeneracy order: full counts at
    k = 3..8 equal the oracle's for node and edge tasks; at k = 9..11 edge-task
    samples equal the oracle's over the same tasks, and at k = 9 the full node
    and edge counts agree (no overflow: C(218, 11) bounds every task)."""
    G = K.Graph.generate(K.GEN_RMAT, seed=3, scale=16, **RMAT)
    Og, orp, oc = _oriented_pair(G, "degen")
    assert Og.info()["max_core"] == 217
    for k in range(3, 12):
        edge = Og.count(k, "pivot", "edge")
        assert not edge.overflow, k
        if k <= 9:  # (node tasks at k = 10, 11: 24 s and 141 s; edge tasks 3 s and 24 s)
            node = Og.count(k, "pivot", "node")
            assert node.count == edge.count and not node.overflow, k
        _log(f"k={k}: edge {edge.count} ({edge.raw['count_ms']:.1f} ms)")
        t = time.perf_counter()
        if k <= 8:
            exp = O.count(orp, oc, k, "pivot_tails", "edge", threads=NT)
            assert edge.count == exp["count"], k
        else:
            stride = {9: 31, 10: 211, 11: 2003}[k]
            exp = O.count(orp, oc, k, "pivot_tails", "edge", stride=stride, offset=5, threads=NT)
            r = Og.count_sampled(k, stride, 5, "pivot", "edge")
            assert r.count == exp["count"] and r.raw["tasks"] == exp["tasks"], k
        _log(f"  oracle {exp['count']} ({exp['tasks']} tasks, {time.perf_counter() - t:.1f}s)")
