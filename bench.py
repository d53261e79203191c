"""Benchmark: exact k-clique counting on the 11.1M-edge Chung-Lu power-law graph.

Workload (one step): the oriented Chung-Lu graph n=1,700,000, m=11,100,000,
degree exponent 2.1, max expected degree 35,000, seed 2, degeneracy
orientation, edge-parallel pivoting with bitmap encoding, counted at
k = 5, 6, 7, 8.  The graph is generated and oriented once, outside the timed
region, and stays resident in HBM.  ``value`` is the reference's ``teps``
aggregated over the step: 4 * m_oriented / (wall time of the four counts).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--ks 5,6,7,8] [--no-cpu]

Prints one JSON line.  The roofline object describes the dominant kernel (the
k=8 launch of the largest-time size class), timed with HIP events on the
library's stream; the cpu_baseline leg times the CPU port of the same
algorithm (oracle/, closed-form SCT tails) on a strided sample of edge tasks.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "k-cliques/s and edges/s at k=5..8; serving p50/p99 vs per-call plain kernels"
HBM_PEAK_GBS = 8000.0
CL = dict(n=1_700_000, m=11_100_000, gamma=2.1, max_degree=35000.0)
SEED = 2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def run_batch(args):
    """Batched small-graph mix: the per-call plain path (one kc_count per
    graph) against one kc_batch_count over all graphs; same counts required."""
    from concurrent.futures import ThreadPoolExecutor
    import kcgpu_amd as K
    specs = batch_mix(args.graphs)

    def make(sp):
        if sp["kind"] == "ba":
            G = K.Graph.generate(K.GEN_BA, seed=sp["seed"], n=sp["n"], attach=sp["attach"])
        else:
            G = K.Graph.generate(K.GEN_ER, seed=sp["seed"], n=sp["n"], m=sp["m"])
        return G, G.autotune(sp["k"])

    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        made = list(ex.map(make, specs))
    log(f"generated {len(made)} graphs in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    ors = [G.orient(plan["orient"]) for G, plan in made]
    log(f"oriented in {time.perf_counter() - t:.1f}s")
    m_total = sum(o.info()["m"] for o in ors)
    queries = [dict(k=sp["k"], algo=plan["algo"], gran=plan["gran"], part_size=plan["part_size"])
               for sp, (G, plan) in zip(specs, made)]
    # warm both paths once
    ors[0].count(**queries[0])
    K.batch_count([(ors[0], queries[0])])
    t = time.perf_counter()
    plain = [o.count(**q).count for o, q in zip(ors, queries)]
    t_plain = time.perf_counter() - t
    t = time.perf_counter()
    res, st = K.batch_count(list(zip(ors, queries)))
    t_batch = time.perf_counter() - t
    same = [r.count for r in res] == plain
    log(f"plain {t_plain:.3f}s batched {t_batch:.3f}s identical={same}")
    if not same:
        raise SystemExit("batched counts differ from the per-call counts")
    out = {"metric": "batched small-graph mix: queries/s, one persistent launch vs per-call plain kernels",
           "value": len(ors) / t_batch, "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "u64", "data": "synthetic",
           "config": {"workload": f"{len(ors)} graphs, half BA n~U[200,20000] attach~U[3,20], half ER "
                                  f"n~U[200,20000] avg degree~U[4,40], seed(i)=splitmix64(4^i), k~U{{4..7}}, autotuned",
                      "m_oriented_total": m_total, "plain_s": t_plain, "batched_s": t_batch,
                      "speedup_vs_plain": t_plain / t_batch, "plain_queries_per_s": len(ors) / t_plain,
                      "batched_edges_per_s": m_total / t_batch, "device_ms": st["total_ms"],
                      "tasks": st["tasks"],
                      "plans": {x: sum(1 for _, p in made if p["orient"] == x) for x in ("degree", "degen")}}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--ks", default="5,6,7,8")
    ap.add_argument("--part-size", type=int, default=0, help="0 = autotune")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=20000)
    ap.add_argument("--workload", choices=["c2", "batch"], default="c2",
                    help="c2: the headline k=5..8 count (default); batch: the 4096-graph batched mix")
    ap.add_argument("--graphs", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X only: this build has no multi-GPU path "
              "(tasks shard by edge with a final 8-byte sum; see DESIGN.md). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    ks = [int(x) for x in args.ks.split(",")]
    if args.workload == "batch":
        return run_batch(args)

    import numpy as np
    import kcgpu_amd as K

    dev = K.device_info()
    log("device", dev)
    t0 = time.perf_counter()
    G = K.Graph.generate(K.GEN_CHUNG_LU, seed=SEED, **CL)
    ginfo = G.info()
    log("generated", ginfo, f"{time.perf_counter() - t0:.1f}s")
    Og = G.orient("degen")
    oinfo = Og.info()
    log("oriented", oinfo)
    m_or = oinfo["m"]
    auto = args.part_size == 0
    P = args.part_size or 1

    def step():
        res = {}
        for k in ks:
            r = Og.count(k, "pivot", "edge", part_size=P, auto_tune=auto)
            res[k] = r
        return res

    for i in range(args.warmup):
        log(f"warmup {i}")
        step()
    times, dev_ms, last = [], {k: [] for k in ks}, None
    dom = {"ms": [], "bytes": [], "class": None}
    t_all = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        last = step()
        times.append(time.perf_counter() - t)
        for k in ks:
            dev_ms[k].append(last[k].raw["count_ms"])
        kd = max(ks)
        dom["ms"].append(last[kd].raw["dominant_ms"])
        dom["bytes"].append(last[kd].raw["dominant_bytes"])
        dom["class"] = last[kd].raw["dominant_class"]
        log(f"step {i}: {times[-1]:.3f}s", {k: last[k].count for k in ks})
    wall = time.perf_counter() - t_all
    per_step = wall / args.steps
    value = len(ks) * m_or / per_step

    per_k = {}
    for k in ks:
        ms = float(np.mean(dev_ms[k]))
        c = last[k].count
        per_k[str(k)] = {"count": c, "device_ms": round(ms, 3), "kcliques_per_s": c / (ms / 1e3) if ms else None,
                         "edges_per_s": m_or / (ms / 1e3) if ms else None, "tasks": last[k].raw["tasks"],
                         "tree_nodes": last[k].raw["tree_nodes"], "part_size": last[k].raw["part_size"]}

    avg_ms = float(np.mean(dom["ms"]))
    avg_bytes = float(np.mean(dom["bytes"]))
    achieved = avg_bytes / (avg_ms / 1e3) / 1e9 if avg_ms else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": f"k_count<pivot,edge> size class {dom['class']} at k={max(ks)}",
                "kernel_ms": round(avg_ms, 3), "alg_bytes": int(avg_bytes)}

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        orp, oc, _ = Og.export()
        nt = min(16, os.cpu_count() or 1)
        stride = args.cpu_stride
        t = time.perf_counter()
        tasks = 0
        for k in ks:
            r = O.count(orp, oc, k, "pivot_tails", "edge", stride=stride, offset=0, threads=nt)
            tasks += r["tasks"]
        tc = time.perf_counter() - t
        sample_edges = (m_or + stride - 1) // stride
        cpu = {"value": len(ks) * sample_edges / tc, "unit": "edges/s", "cores": nt, "kind": "port",
               "sample": f"every {stride}th oriented edge ({sample_edges} edges, {tasks} tasks with |T|>=k-2) "
                         f"at k={','.join(map(str, ks))}; {tc:.1f}s on {nt} OpenMP threads",
               "seconds": round(tc, 2)}
        log("cpu baseline", cpu)

    out = {"metric": METRIC, "value": value, "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": per_step * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u64", "data": "synthetic",
           "config": {"workload": "chung-lu n=1.7M m=11.1M gamma=2.1 maxdeg~35000 seed=2; "
                                  "-o degen -p edge -q p<P>b; k=" + ",".join(map(str, ks)),
                      "n": ginfo["n"], "m": ginfo["m"], "m_oriented": m_or, "max_core": oinfo["max_core"],
                      "max_out_degree": oinfo["max_out_degree"], "checksum": ginfo["checksum"],
                      "orient_ms": round(oinfo["preprocess_ms"], 3), "part_size": "auto" if auto else P,
                      "per_k": per_k},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
