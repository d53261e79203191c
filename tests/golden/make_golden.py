"""Writes tests/golden/cliques.json: small graphs with their k-clique counts.

Counts (total, and per vertex for k = 3..6) come from networkx (``enumerate_all_cliques``), an implementation
independent of both the oracle and the HIP kernels, run in the build
container only; the JSON is what the tests read.  Closed-form families are
added with their formulas so that a mistake in networkx use would also show.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
from math import comb

import networkx as nx

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cliques.json")
KMAX = 10


def counts(G):
    c = {}
    for cl in nx.enumerate_all_cliques(G):
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
    with open(OUT, "w") as f:
        json.dump({"kmax": KMAX, "graphs": g}, f, separators=(",", ":"))
    print("wrote", OUT, len(g), "graphs")


if __name__ == "__main__":
    main()
