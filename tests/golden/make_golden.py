"""Writes tests/golden/cliques.json: small graphs with their k-clique counts.

Counts (total, and per vertex for k = 3..6) come from networkx (``enumerate_all_cliques``), an implementation
independent of both the oracle and the HIP kernels, run once on a host
with networkx; the JSON is what the tests read.  Closed-form families are
added with their formulas so that a mistake in networkx use would also show.

"near_cliques" are larger graphs (200 to 1100 vertices) given by their
complement H: G = complement(H), where H is a disjoint union of small
components (triangles, K4, paths, stars, C4, paws, diamonds, random graphs
on 4..7 vertices) and isolated vertices.  The k-cliques of G are the
independent k-sets of H, and the independence polynomial of a disjoint
union is the product of its components' polynomials, so the exact counts
(Python integers, also above 2^64) follow from brute force over each
component.  The formula is checked against networkx on small instances.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
from math import comb

import networkx as nx

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cliques.json")
KMAX = 11


def counts(G):
    c = {}
    for cl in nx.enumerate_all_cliques(G):  # cliques come in nondecreasing size
        if len(cl) > KMAX:
            break
        c[len(cl)] = c.get(len(cl), 0) + 1
    return {k: c.get(k, 0) for k in range(1, KMAX + 1)}


LOCAL_K = (3, 4, 5, 6)


def local_counts(G):
    """per-vertex k-clique counts for k in LOCAL_K (``-m lkc``)"""
    loc = {k: [0] * G.number_of_nodes() for k in LOCAL_K}
    for cl in nx.enumerate_all_cliques(G):
        if len(cl) in loc:
            for v in cl:
                loc[len(cl)][v] += 1
        if len(cl) > max(LOCAL_K):
            break
    return loc


def entry(name, G, closed=None):
    G = nx.convert_node_labels_to_integers(G)
    e = {
        "name": name,
        "n": G.number_of_nodes(),
        "edges": sorted([min(u, v), max(u, v)] for u, v in G.edges()),
        "counts": counts(G),
        "max_core": max(nx.core_number(G).values()) if G.number_of_edges() else 0,
        "local": local_counts(G),
    }
    if closed is not None:
        assert all(e["counts"][k] == closed(k) for k in range(1, KMAX + 1)), name
        e["closed_form"] = True
    return e


def planted(n, p, sizes, seed):
    G = nx.gnp_random_graph(n, p, seed=seed)
    import random
    r = random.Random(seed)
    for s in sizes:
        vs = r.sample(range(n), s)
        for i in range(s):
            for j in range(i + 1, s):
                G.add_edge(vs[i], vs[j])
    return G


# ---- near-cliques with exact counts from the complement's independence polynomial ----
SHAPES = {
    "K2": [(0, 1)], "P3": [(0, 1), (1, 2)], "K3": [(0, 1), (0, 2), (1, 2)],
    "P4": [(0, 1), (1, 2), (2, 3)], "S3": [(0, 1), (0, 2), (0, 3)], "C4": [(0, 1), (1, 2), (2, 3), (0, 3)],
    "paw": [(0, 1), (0, 2), (1, 2), (2, 3)], "diamond": [(0, 1), (0, 2), (1, 2), (1, 3), (2, 3)],
    "K4": [(a, b) for a in range(4) for b in range(a + 1, 4)],
}
ALL_SHAPES = ["K3", "K4", "P4", "S3", "C4", "paw", "diamond", "P3", "K2"]


def independent_set_counts(s, edges, kmax):
    """number of independent sets of each size 0..kmax of a graph on s <= 64 vertices"""
    adj = [0] * s
    for a, b in edges:
        adj[a] |= 1 << b
        adj[b] |= 1 << a
    cnt = [0] * (kmax + 1)

    def rec(cand, size):
        cnt[size] += 1
        if size == kmax:
            return
        while cand:
            low = cand & -cand
            cand ^= low
            rec(cand & ~adj[low.bit_length() - 1], size + 1)

    rec((1 << s) - 1, 0)
    return cnt


def polymul(a, b, kmax):
    out = [0] * (kmax + 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b[:kmax + 1 - i]):
            out[i + j] += x * y
    return out


def near_clique(name, n, shapes, n_random, seed, kmax=KMAX):
    """G = complement of H; H = the listed shapes, n_random random graphs on
    4..7 vertices (edge probability 1/2) and isolated vertices, placed on a
    seeded random permutation of the n vertices."""
    import random
    r = random.Random(seed)
    perm = list(range(n))
    r.shuffle(perm)
    comps = [SHAPES[x] for x in shapes]
    for _ in range(n_random):
        s = r.randint(4, 7)
        comps.append([(a, b) for a in range(s) for b in range(a + 1, s) if r.random() < 0.5])
    poly = [1] + [0] * kmax
    h_edges, i = [], 0
    for es in comps:
        s = 1 + max((max(e) for e in es), default=0)
        poly = polymul(poly, independent_set_counts(s, es, kmax), kmax)
        h_edges += [sorted((perm[i + a], perm[i + b])) for a, b in es]
        i += s
    assert i <= n
    for _ in range(n - i):  # isolated vertices of H: x + 1 each
        poly = polymul(poly, [1, 1] + [0] * (kmax - 1), kmax)
    return {"name": name, "n": n, "h_edges": sorted(h_edges), "counts": {k: poly[k] for k in range(1, kmax + 1)}}


def check_near_clique_formula():
    """the product formula against networkx clique counts on small instances"""
    for seed in range(4):
        e = near_clique("t", 18, [ALL_SHAPES[seed], ALL_SHAPES[seed + 4]], 1, seed)
        H = nx.Graph()
        H.add_nodes_from(range(e["n"]))
        H.add_edges_from(map(tuple, e["h_edges"]))
        assert counts(nx.complement(H)) == e["counts"], seed


def main():
    g = []
    g.append(entry("K12", nx.complete_graph(12), lambda k: comb(12, k)))
    g.append(entry("K25", nx.complete_graph(25), lambda k: comb(25, k)))
    g.append(entry("turan_24_4", nx.turan_graph(24, 4), lambda k: comb(4, k) * 6 ** k))
    g.append(entry("disjoint_K5_K7_K9", nx.disjoint_union_all([nx.complete_graph(a) for a in (5, 7, 9)]),
                   lambda k: comb(5, k) + comb(7, k) + comb(9, k)))
    g.append(entry("petersen", nx.petersen_graph()))
    g.append(entry("cycle_50", nx.cycle_graph(50)))
    g.append(entry("star_40", nx.star_graph(40)))
    g.append(entry("moon_moser_15", nx.complement(nx.disjoint_union_all([nx.complete_graph(3)] * 5)),
                   lambda k: comb(5, k) * 3 ** k))
    for i, (n, p) in enumerate([(40, 0.5), (60, 0.3), (90, 0.2), (64, 0.6), (120, 0.15)]):
        g.append(entry(f"gnp_{n}_{p}_s{i}", nx.gnp_random_graph(n, p, seed=100 + i)))
    g.append(entry("ba_300_6", nx.barabasi_albert_graph(300, 6, seed=7)))
    g.append(entry("ba_500_10", nx.barabasi_albert_graph(500, 10, seed=8)))
    g.append(entry("planted_200", planted(200, 0.05, [14, 11, 9], seed=9)))
    g.append(entry("planted_400", planted(400, 0.02, [20, 16, 12, 8], seed=10)))
    g.append(entry("ws_200", nx.watts_strogatz_graph(200, 12, 0.2, seed=11)))
    g.append(entry("gnp_150_0.4", nx.gnp_random_graph(150, 0.4, seed=12)))
    check_near_clique_formula()
    nc = [
        # tasks of up to 199 vertices: the register size classes
        near_clique("nc200_a", 200, ALL_SHAPES * 2, 0, 21),
        near_clique("nc200_b", 200, ALL_SHAPES, 6, 22),
        # up to 299: the wide kernel, bitmap in LDS
        near_clique("nc300", 300, ALL_SHAPES * 2, 2, 23),
        # up to 1099: the wide kernel's HBM slice
        near_clique("nc1100", 1100, ALL_SHAPES * 2, 4, 24),
    ]
    with open(OUT, "w") as f:
        json.dump({"kmax": KMAX, "graphs": g, "near_cliques": nc}, f, separators=(",", ":"))
    print("wrote", OUT, len(g), "graphs", len(nc), "near-cliques")


if __name__ == "__main__":
    main()
