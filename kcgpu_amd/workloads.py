"""Seeded synthetic workloads shared by bench.py and the tests: the batched
small-graph mix (4,096 graphs of uneven size) and the serving mix (resident
graphs of up to 2 M edges).  Graphs come from the library's own generators
(kc_generate), so the same bytes are produced on every host."""
from __future__ import annotations


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def batch_mix(count=4096):
    """The batched workload: graph i has seed splitmix64(4 ^ i); even i are
    Barabasi-Albert (n ~ U[200, 20000], attach ~ U[3, 20]), odd i Erdos-Renyi
    (n ~ U[200, 20000], average degree ~ U[4, 40]); k ~ U{4..7}.  Parameters
    are drawn from numpy's PCG64 seeded with the graph seed."""
    import numpy as np
    out = []
    for i in range(count):
        s = splitmix64(4 ^ i)
        r = np.random.default_rng(s)
        n = int(r.integers(200, 20001))
        if i % 2 == 0:
            spec = dict(kind="ba", n=n, attach=int(r.integers(3, 21)))
        else:
            spec = dict(kind="er", n=n, m=n * int(r.integers(4, 41)) // 2)
        spec["k"] = int(r.integers(4, 8))
        spec["seed"] = s
        out.append(spec)
    return out


def serve_mix(count, seed):
    """Resident graphs for the serving workload: a seeded mix of the ER,
    Chung-Lu, R-MAT and BA generators, each at most 2M edges."""
    import numpy as np
    r = np.random.default_rng(seed)
    out = []
    for i in range(count):
        t = i % 4
        if t == 0:
            n = int(r.integers(5000, 100001))
            out.append(dict(kind="er", n=n, m=min(2_000_000, n * int(r.integers(4, 41)) // 2)))
        elif t == 1:
            n = int(r.integers(20000, 300001))
            out.append(dict(kind="cl", n=n, m=min(2_000_000, n * int(r.integers(3, 7))), gamma=2.1,
                            max_degree=float(r.integers(300, 2001))))
        elif t == 2:
            out.append(dict(kind="rmat", scale=int(r.integers(11, 16)), edge_factor=16))
        else:
            n = int(r.integers(1000, 50001))
            out.append(dict(kind="ba", n=n, attach=int(r.integers(3, 21))))
        out[-1]["seed"] = int(r.integers(1, 2**62))
    return out


def make_graph(K, sp):
    """kcgpu_amd.Graph of a workload spec (batch_mix / serve_mix entry)"""
    if sp["kind"] == "er":
        return K.Graph.generate(K.GEN_ER, seed=sp["seed"], n=sp["n"], m=sp["m"])
    if sp["kind"] == "cl":
        return K.Graph.generate(K.GEN_CHUNG_LU, seed=sp["seed"], n=sp["n"], m=sp["m"], gamma=sp["gamma"],
                                max_degree=sp["max_degree"])
    if sp["kind"] == "rmat":
        return K.Graph.generate(K.GEN_RMAT, seed=sp["seed"], scale=sp["scale"], edge_factor=sp["edge_factor"],
                                a=0.57, b=0.19, c=0.19)
    return K.Graph.generate(K.GEN_BA, seed=sp["seed"], n=sp["n"], attach=sp["attach"])
