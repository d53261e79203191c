"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Every oracle function is checked against something it does not share code
with: networkx-generated golden counts and closed forms (tests/golden),
brute force over all k-subsets, the synchronous-peeling definition of the
degeneracy key recomputed from scratch, hand-worked examples, and exact
binomials from Python integers.
"""
from math import comb

import numpy as np
import pytest

from oracle import oracle as O


def _csr(g):
    e = np.array(g["edges"], dtype=np.int64).reshape(-1, 2)
    return O.csr_from_edges(e[:, 0], e[:, 1], g["n"])


# ---- CSR builder -------------------------------------------------------------
def test_csr_builder_cleans_edges():
    # duplicates, both directions, self-loops; a self-loop does not create a
    # row (include/CSRCOO.cuh:51-52 skips it before extending rowPtr)
    src = [0, 1, 1, 2, 2, 3, 5]
    dst = [1, 0, 2, 1, 2, 0, 5]
    rp, col = O.csr_from_edges(src, dst)
    assert rp.tolist() == [0, 2, 4, 5, 6]
    assert col.tolist() == [1, 3, 0, 2, 1, 0]


# ---- orientation -------------------------------------------------------------
def test_orient_rule_and_halving(golden):
    for g in golden:
        rp, col = _csr(g)
        deg = np.diff(rp).astype(np.int64)
        orp, oc = O.orient(rp, col)
        assert oc.size * 2 == col.size
        src = np.repeat(np.arange(g["n"]), np.diff(orp).astype(np.int64))
        key = list(zip(deg[src], src))
        keyv = list(zip(deg[oc.astype(np.int64)], oc.astype(np.int64)))
        assert all(a < b for a, b in zip(key, keyv))
        # every undirected edge exactly once
        und = {(min(a, b), max(a, b)) for a, b in zip(src.tolist(), oc.tolist())}
        assert und == {tuple(e) for e in g["edges"]}
        # the C restatement agrees with the numpy one, for both keys
        orp2, oc2 = O.orient_c(rp, col)
        assert (orp == orp2).all() and (oc == oc2).all()
        pr, _ = O.peel_priority(rp, col)
        a, b = O.orient(rp, col, pr), O.orient_c(rp, col, pr)
        assert (a[0] == b[0]).all() and (a[1] == b[1]).all()


# ---- degeneracy peel ---------------------------------------------------------
def test_peel_hand_examples():
    # triangle 0-1-2, pendant 3 on 2, isolated 4: level 0 peels 4 (round 0),
    # level 1 peels 3 (round 1); the level-2 round empties the graph and keeps nnz
    rp, col = O.csr_from_edges([0, 1, 0, 2], [1, 2, 2, 3], 5)
    pr, mc = O.peel_priority(rp, col)
    assert pr.tolist() == [8, 8, 8, 1, 0] and mc == 2
    # path 0-1-2-3-4: rounds {0,4}, {1,3}; vertex 2 is the final round
    rp, col = O.csr_from_edges([0, 1, 2, 3], [1, 2, 3, 4], 5)
    pr, mc = O.peel_priority(rp, col)
    assert pr.tolist() == [0, 1, 8, 1, 0] and mc == 1


def _peel_by_definition(rp, col):
    """Synchronous peeling from its definition: round r removes every live
    vertex whose live degree is <= L(r), L(r) = max(L(r-1), min live degree);
    the round that removes the last vertices is not numbered."""
    n = rp.size - 1
    nnz = int(rp[-1])
    adj = [col[rp[v]:rp[v + 1]].tolist() for v in range(n)]
    live = set(range(n))
    prio = [nnz] * n
    level, r = 0, 0
    while live:
        d = {v: sum(1 for x in adj[v] if x in live) for v in live}
        level = max(level, min(d.values()))
        s = [v for v in live if d[v] <= level]
        if len(s) == len(live):
            return prio, level
        for v in s:
            prio[v] = r
        live -= set(s)
        r += 1
    return prio, level


def test_peel_matches_definition_and_core_number(golden):
    for g in golden:
        rp, col = _csr(g)
        pr, mc = O.peel_priority(rp, col)
        ref, lev = _peel_by_definition(rp, col)
        assert pr.tolist() == ref, g["name"]
        assert mc == lev == g["max_core"], g["name"]
        orp, _ = O.orient(rp, col, pr)
        assert int(np.diff(orp).max(initial=0)) <= mc


# ---- binomials -----------------------------------------------------------------
def test_binom_exact_and_overflow():
    for n in range(0, 120, 7):
        for r in range(0, 40):
            v, ovf = O.binom(n, r)
            exact = comb(n, r)
            assert ovf == (exact >= 2 ** 64)
            if not ovf:
                assert v == exact
    assert O.binom(67, 33) == (comb(67, 33), False)  # 1.4e19 < 2^64
    assert O.binom(68, 34)[1]                          # 2.8e19 > 2^64


# ---- counting -----------------------------------------------------------------------
@pytest.mark.parametrize("algo", ["orient", "pivot", "pivot_tails"])
@pytest.mark.parametrize("gran", ["node", "edge"])
def test_counts_vs_golden(golden, algo, gran):
    for g in golden:
        rp, col = _csr(g)
        pr, _ = O.peel_priority(rp, col)
        for key in (None, pr):
            orp, oc = O.orient(rp, col, key)
            for k in range(1, 11):
                if algo == "orient" and g["name"] == "K25" and k > 8:
                    continue  # orientation DFS on K25 enumerates 2e6 cliques per k: covered by pivot
                r = O.count(orp, oc, k, algo, gran)
                assert r["count"] == g["counts"][k] and not r["overflow"], (g["name"], k, algo, gran)


def test_counts_k11_vs_golden(golden):
    """k = 11 (the top of the counted range) from networkx"""
    for g in golden:
        rp, col = _csr(g)
        orp, oc = O.orient(rp, col, O.peel_priority(rp, col)[0])
        for algo in ("pivot", "pivot_tails"):
            for gran in ("node", "edge"):
                r = O.count(orp, oc, 11, algo, gran)
                assert r["count"] == g["counts"][11] and not r["overflow"], (g["name"], algo, gran)


def test_near_clique_closed_forms(near_cliques):
    """The oracle on 200-vertex complements of sparse graphs at k = 8..11:
    equal to the counts from the complement's independence polynomial."""
    from conftest import complement_edges
    for g in near_cliques:
        if g["n"] > 200:
            continue
        a, b = complement_edges(g["n"], g["h_edges"])
        rp, col = O.csr_from_edges(a, b, g["n"])
        for key in (None, O.peel_priority(rp, col)[0]):
            orp, oc = O.orient(rp, col, key)
            for k in (8, 9, 10, 11):
                r = O.count(orp, oc, k, "pivot_tails", "node" if k % 2 else "edge")
                assert r["count"] == g["counts"][k] and not r["overflow"], (g["name"], k)
    # a planted mistake: one complement edge left out changes the count
    g = near_cliques[0]
    a, b = complement_edges(g["n"], g["h_edges"][1:])
    rp, col = O.csr_from_edges(a, b, g["n"])
    orp, oc = O.orient(rp, col)
    assert O.count(orp, oc, 8, "pivot_tails", "node")["count"] != g["counts"][8]


def test_brute_force_vs_golden(golden):
    for g in golden:
        if g["n"] > 64:
            continue
        e = np.array(g["edges"], dtype=np.int64).reshape(-1, 2)
        for k in range(1, 9):
            v, o = O.brute_force(g["n"], e[:, 0], e[:, 1], k)
            assert v == g["counts"][k] and not o, (g["name"], k)


def test_closed_forms_present(golden):
    names = {g["name"] for g in golden if g.get("closed_form")}
    assert {"K12", "K25", "turan_24_4", "disjoint_K5_K7_K9", "moon_moser_15"} <= names


def test_pivot_overflow_detection():
    # K70: C(70, 30) > 2^64 must be flagged; C(70, 10) must be exact
    n = 70
    src, dst = zip(*[(i, j) for i in range(n) for j in range(i + 1, n)])
    rp, col = O.csr_from_edges(src, dst, n)
    orp, oc = O.orient(rp, col)
    for gran in ("node", "edge"):
        r = O.count(orp, oc, 30, "pivot", gran)
        assert r["overflow"]
        r = O.count(orp, oc, 10, "pivot", gran)
        assert r["count"] == comb(70, 10) and not r["overflow"]


def test_sampled_tasks_partition_the_count(golden):
    g = next(x for x in golden if x["name"] == "planted_400")
    rp, col = _csr(g)
    orp, oc = O.orient(rp, col)
    for algo in ("orient", "pivot", "pivot_tails"):
        for gran in ("node", "edge"):
            parts = [O.count(orp, oc, 6, algo, gran, stride=3, offset=o)["count"] for o in range(3)]
            assert sum(parts) == g["counts"][6]


def test_planted_mistake_is_caught(golden):
    """A wrong orientation (keeping edges in both directions) double counts:
    the golden comparison used above must reject it."""
    g = next(x for x in golden if x["name"] == "gnp_40_0.5_s0")
    rp, col = _csr(g)
    r = O.count(rp.astype(np.uint64), col, 4, "orient", "node")  # unoriented CSR
    assert r["count"] != g["counts"][4]


# ---- local (per-vertex) counts, -m lkc ------------------------------------------
@pytest.mark.parametrize("algo", ["orient", "pivot", "pivot_tails"])
@pytest.mark.parametrize("gran", ["node", "edge"])
def test_local_counts_vs_golden(golden, algo, gran):
    for g in golden:
        if algo == "orient" and g["name"] == "K25":
            continue
        rp, col = _csr(g)
        pr, _ = O.peel_priority(rp, col)
        for key in (None, pr):
            orp, oc = O.orient(rp, col, key)
            for k in (3, 4, 5, 6):
                loc, ovf = O.count_local(orp, oc, k, algo, gran)
                assert not ovf
                assert loc.tolist() == g["local"][str(k)], (g["name"], k, algo, gran)


def test_local_counts_closed_forms_and_invariants(golden):
    # K_n: every vertex lies in C(n-1, k-1) k-cliques; k=1 -> 1; k=2 -> degree
    n = 30
    src, dst = zip(*[(i, j) for i in range(n) for j in range(i + 1, n)])
    rp, col = O.csr_from_edges(src, dst, n)
    orp, oc = O.orient(rp, col)
    for k in (1, 2, 3, 7, 12):
        for algo in ("pivot", "pivot_tails"):
            for gran in ("node", "edge"):
                loc, ovf = O.count_local(orp, oc, k, algo, gran)
                assert (loc == comb(n - 1, k - 1)).all() and not ovf, (k, algo, gran)
    # sum over vertices = k * count, for counts beyond the golden k range
    g = next(x for x in golden if x["name"] == "gnp_150_0.4")
    rp, col = _csr(g)
    pr, _ = O.peel_priority(rp, col)
    orp, oc = O.orient(rp, col, pr)
    for k in (7, 8, 9):
        for gran in ("node", "edge"):
            loc, _ = O.count_local(orp, oc, k, "pivot_tails", gran)
            assert int(loc.sum(dtype=np.uint64)) == k * g["counts"][k]
            assert (loc == O.count_local(orp, oc, k, "pivot", gran)[0]).all()
    # overflow of one vertex's count is flagged: K70, k = 31 -> C(69, 30) > 2^64
    n = 70
    src, dst = zip(*[(i, j) for i in range(n) for j in range(i + 1, n)])
    rp, col = O.csr_from_edges(src, dst, n)
    orp, oc = O.orient(rp, col)
    assert O.count_local(orp, oc, 31, "pivot_tails", "node")[1]
    loc, ovf = O.count_local(orp, oc, 10, "pivot_tails", "edge")
    assert (loc == comb(69, 9)).all() and not ovf

