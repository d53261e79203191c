"""The CSR built on the device (kc_graph_from_edges_device, kc_orient_edges)
against the host builder (kc_graph_from_edges), entry for entry, and counts on
the device-built orientation against the oracle."""
import time

import numpy as np
import pytest

import kcgpu_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _same_csr(a, b):
    ra, ca = a.csr()
    rb, cb = b.csr()
    assert np.array_equal(ra, rb) and np.array_equal(ca, cb)
    assert a.info() == b.info()


def _messy(rng, n, m):
    """random edges with loops, repeats and both directions of some edges"""
    s = rng.integers(0, n, m, dtype=np.uint32)
    d = rng.integers(0, n, m, dtype=np.uint32)
    k = m // 10
    s = np.concatenate([s, d[:k], s[:k], np.arange(min(n, 50), dtype=np.uint32)])
    d = np.concatenate([d, s[:k], d[:k], np.arange(min(n, 50), dtype=np.uint32)])
    p = rng.permutation(s.size)
    return s[p], d[p]


@pytest.mark.parametrize("n,m", [(1, 0), (2, 1), (10, 40), (1000, 5000), (50000, 400000)])
def test_device_csr_equals_host_random(n, m):
    rng = np.random.default_rng(n + m)
    s, d = _messy(rng, n, m)
    _same_csr(K.Graph.from_edges_device(s, d), K.Graph.from_edges(s, d))
    _same_csr(K.Graph.from_edges_device(s, d, n=n + 7), K.Graph.from_edges(s, d, n=n + 7))


def test_device_csr_hubs_and_edge_cases():
    rng = np.random.default_rng(3)
    # hubs: rows of 3,000 to 60,000 entries (the bitmap path), with repeats
    hubs = []
    for h, deg in ((0, 3000), (7, 20000), (123, 60000)):
        leaves = rng.integers(0, 100000, deg, dtype=np.uint32)
        hubs.append((np.full(deg, h, np.uint32), leaves))
    s = np.concatenate([x for x, _ in hubs] + [rng.integers(0, 100000, 200000, dtype=np.uint32)])
    d = np.concatenate([y for _, y in hubs] + [rng.integers(0, 100000, 200000, dtype=np.uint32)])
    Gd, Gh = K.Graph.from_edges_device(s, d), K.Graph.from_edges(s, d)
    _same_csr(Gd, Gh)
    assert Gd.info()["max_degree"] > 2048
    # only loops, and nothing at all
    loops = np.arange(10, dtype=np.uint32)
    _same_csr(K.Graph.from_edges_device(loops, loops), K.Graph.from_edges(loops, loops))
    empty = np.zeros(0, np.uint32)
    _same_csr(K.Graph.from_edges_device(empty, empty), K.Graph.from_edges(empty, empty))
    # an id at or above an explicit n is an error, as on the host
    with pytest.raises(K.KCError) as e:
        K.Graph.from_edges_device(np.array([1], np.uint32), np.array([5], np.uint32), n=5)
    assert e.value.code == K.KC_ERR_ARG


@pytest.mark.parametrize("mode", ["degree", "degen"])
def test_orient_edges_equals_host_path_and_oracle(golden, mode):
    # the oriented CSR of the device-built graph equals the host path's, and
    # its counts equal the golden counts
    for g in golden:
        e = np.array(g["edges"], dtype=np.uint32).reshape(-1, 2)
        Og = K.orient_edges(e[:, 0], e[:, 1], mode, n=g["n"])
        Oh = K.Graph.from_edges(e[:, 0], e[:, 1], g["n"]).orient(mode)
        for x, y in zip(Og.export(), Oh.export()):
            assert np.array_equal(x, y), g["name"]
        for k in (3, 4, 5):
            if k in g["counts"]:
                assert Og.count(k, "pivot", "edge").count == g["counts"][k], (g["name"], k)


def test_orient_edges_rmat_scale16_and_full_size_csr():
    # R-MAT scale 16 edges in shuffled order with repeats: orientation and
    # counts through the device builder equal the host path's and the oracle's
    G = K.Graph.generate(K.GEN_RMAT, seed=3, scale=16, edge_factor=16, a=0.57, b=0.19, c=0.19)
    rp, col = G.csr()
    src = np.repeat(np.arange(rp.size - 1, dtype=np.uint32), np.diff(rp).astype(np.int64))
    rng = np.random.default_rng(5)
    p = rng.permutation(src.size)
    s, d = src[p], col[p]  # every edge twice (both directions), in random order
    n = rp.size - 1  # (the highest ids of the generator's range may be isolated)
    for mode in ("degree", "degen"):
        Og = K.orient_edges(s, d, mode, n=n)
        Oh = G.orient(mode)
        for x, y in zip(Og.export(), Oh.export()):
            assert np.array_equal(x, y), mode
    Od = K.orient_edges(s, d, "degen", n=n)
    orp, oc, _ = Od.export()
    assert Od.count(6, "pivot", "edge").count == \
        O.count(orp.astype(np.uint64), oc, 6, "pivot_tails", "edge", threads=8)["count"]
    # the R-MAT scale 22 graph (64.2 M edges, hubs of 160,000): one direction
    # of each edge, shuffled; the device-built CSR equals the generator's
    G = K.Graph.generate(K.GEN_RMAT, seed=3, scale=22, edge_factor=16, a=0.57, b=0.19, c=0.19)
    rp, col = G.csr()
    src = np.repeat(np.arange(rp.size - 1, dtype=np.uint32), np.diff(rp).astype(np.int64))
    up = src < col
    s, d = src[up], col[up]
    p = rng.permutation(s.size)
    s, d = s[p], d[p]
    t = time.perf_counter()
    Gd = K.Graph.from_edges_device(s, d, n=rp.size - 1)
    t_dev = time.perf_counter() - t
    t = time.perf_counter()
    Gh = K.Graph.from_edges(s, d, n=rp.size - 1)
    t_host = time.perf_counter() - t
    print(f"R-MAT scale 22 CSR from {s.size} edges: device {t_dev:.2f} s (upload and download included), "
          f"host {t_host:.2f} s", flush=True)
    assert Gd.info() == G.info() == Gh.info()
    rd, cd = Gd.csr()
    assert np.array_equal(rd, rp) and np.array_equal(cd, col)
