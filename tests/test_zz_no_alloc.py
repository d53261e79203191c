"""kc_count_noalloc (no device allocation; the launch scratch is the library's
fixed device array) against kc_count, bit for bit, through the C ABI.

Two calls in a row on the same array, the second on a smaller graph, so a
stale queue head, counter or task list left by the first call would show.
The graph with a planted 1100-vertex clique puts node tasks in the HBM class,
whose bitmap slices come out of the same array.  Edge tasks of a graph whose
scratch exceeds the array must be refused with KC_ERR_LIMIT.
"""
import numpy as np
import pytest

import kcgpu_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _planted(n, m, sizes, seed):
    rng = np.random.default_rng(seed)
    G0 = K.Graph.generate(K.GEN_ER, seed=seed, n=n, m=m)
    rp, col = G0.csr()
    src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    es, ed = [src], [col]
    for s in sizes:
        vs = rng.choice(n, size=s, replace=False).astype(np.uint32)
        a, b = np.triu_indices(s, 1)
        es.append(vs[a])
        ed.append(vs[b])
    return K.Graph.from_edges(np.concatenate(es), np.concatenate(ed), n)


def _same(Og, k, algo, gran):
    a = Og.count(k, algo, gran, part_size=1, check=False)
    b = Og.count_noalloc(k, algo, gran, part_size=1, check=False)
    print(k, algo, gran, a.count, b.count, a.overflow, b.overflow, flush=True)
    assert (b.count, b.overflow) == (a.count, a.overflow), (k, algo, gran)
    assert b.raw["tasks"] == a.raw["tasks"]
    return a


def test_noalloc_equals_count_two_calls_smaller_second():
    big = K.Graph.generate(K.GEN_CHUNG_LU, seed=2, n=60000, m=400000, gamma=2.1, max_degree=3000)
    small = K.Graph.generate(K.GEN_BA, seed=3, n=3000, attach=8)
    Ob, Os = big.orient("degen"), small.orient("degen")
    assert Os.info()["m"] < Ob.info()["m"]
    for k, algo, gran in ((5, "pivot", "edge"), (6, "pivot", "node"), (4, "orient", "edge")):
        _same(Ob, k, algo, gran)   # first call: the larger graph
        r = _same(Os, k, algo, gran)  # second call: smaller, on what the first left behind
        rp, col, _ = Os.export()
        assert r.count == O.count(rp.astype(np.uint64), col, k, "pivot", "node")["count"]


def test_noalloc_hbm_class_and_repeat():
    # out-degrees above 1024: node tasks in the HBM class use bitmap slices
    # carved from the fixed array; then a smaller graph on the same array
    G = _planted(6000, 30000, [300, 1100], seed=5)
    Og = G.orient("degree")
    assert Og.info()["max_out_degree"] > 1024
    for k in (4, 7):
        r = _same(Og, k, "pivot", "node")
        rp, col, _ = Og.export()
        assert r.count == O.count(rp.astype(np.uint64), col, k, "pivot_tails", "node")["count"]
        assert r.raw["class_kernel"][5] != 0
    Os = _planted(2000, 8000, [100], seed=6).orient("degen")
    _same(Os, 5, "pivot", "edge")
    _same(Os, 2, "pivot", "edge")


def test_noalloc_refuses_what_does_not_fit():
    # about 30 M oriented edges: 12 bytes per edge task is more than 256 MiB
    G = K.Graph.generate(K.GEN_RMAT, seed=3, scale=21, edge_factor=16, a=0.57, b=0.19, c=0.19)
    Og = G.orient("degree")
    m = Og.info()["m"]
    assert 12 * m > (K.NOALLOC_SCRATCH_MIB << 20)
    with pytest.raises(K.KCError) as e:
        Og.count_noalloc(4, "pivot", "edge")
    assert e.value.code == K.KC_ERR_LIMIT and "fixed array" in str(e.value)
    # node tasks of the same graph fit
    _same(Og, 4, "pivot", "node")
