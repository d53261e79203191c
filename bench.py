"""Benchmark: exact k-clique counting on the 11.1M-edge Chung-Lu power-law graph.

Workload (one step): the oriented Chung-Lu graph n=1,700,000, m=11,100,000,
degree exponent 2.1, max expected degree 35,000, seed 2, degeneracy
orientation, edge-parallel pivoting with bitmap encoding, counted at
k = 5, 6, 7, 8.  The graph is generated and oriented once, outside the timed
region, and stays resident in HBM.  ``value`` is the reference's ``teps``
aggregated over the step: 4 * m_oriented / (wall time of the four counts).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--ks 5,6,7,8] [--no-cpu]

Prints one JSON line.  The roofline object describes the dominant kernel (the
size class whose tasks hold the device's waves longest over the step),
averaged over its launches, each timed with HIP events on the stream it runs
on; the cpu_baseline leg times the CPU port of the same
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
ORIENT_BOUND = 2e9  # serving: largest clique-enumeration bound an orientation-mode query keeps


def log(*a):
    print(*a, file=sys.stderr, flush=True)


from kcgpu_amd.workloads import batch_mix, make_graph, serve_mix, splitmix64  # noqa: E402,F401


def run_batch(args):
    """Batched small-graph mix: the per-call plain path (one kc_count per
    graph) against one kc_batch_count over all graphs; same counts required."""
    from concurrent.futures import ThreadPoolExecutor
    import kcgpu_amd as K
    specs = batch_mix(args.graphs)

    def make(sp):
        G = make_graph(K, sp)
        return G, G.autotune(sp["k"])

    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        made = list(ex.map(make, specs))
    log(f"generated {len(made)} graphs in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    ors_plain = [G.orient(plan["orient"]) for G, plan in made]
    t_orient_plain = time.perf_counter() - t
    del ors_plain
    t = time.perf_counter()
    ors = K.orient_batch([G for G, _ in made], [plan["orient"] for _, plan in made])
    t_orient_batch = time.perf_counter() - t
    log(f"oriented one by one in {t_orient_plain:.2f}s, batched in {t_orient_batch:.2f}s")
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
                      "orient_plain_s": t_orient_plain, "orient_batched_s": t_orient_batch,
                      "orient_batched_device_ms": ors[0].info()["preprocess_ms"] if ors else 0.0,
                      "batched_edges_per_s": m_total / t_batch, "device_ms": st["total_ms"],
                      "tasks": st["tasks"],
                      "plans": {x: sum(1 for _, p in made if p["orient"] == x) for x in ("degree", "degen")}}}
    print(json.dumps(out), flush=True)


def run_serve(args):
    """Serving: 512 resident graphs + 64 uploaded mid-run, Poisson (graph, k)
    queries (k ~ U{3..10}, orientation or pivot mode, arrival seed 55), HBM
    footprint admission, continuous batching.  Each result is checked against
    the per-call plain path, which is also timed back to back."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import kcgpu_amd as K
    specs = serve_mix(args.resident + args.uploads, seed=5)
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        graphs = list(ex.map(lambda sp: make_graph(K, sp), specs))
    plans = [G.autotune(8) for G in graphs]
    log(f"generated {len(graphs)} graphs in {time.perf_counter() - t:.1f}s")
    # Orientation-mode queries enumerate every clique; the client sends one as
    # pivoting (same count) when the enumeration bound sum_u C(d+(u), k-1) of
    # its graph is above ORIENT_BOUND.  Out-degrees come from one orientation.
    from math import comb
    outdeg_hist = []
    for G, pl in zip(graphs, plans):
        o = G.orient(pl["orient"])
        rp, _, _ = o.export()
        d = np.diff(rp.astype(np.int64))
        outdeg_hist.append(np.bincount(d))
        del o

    def orient_bound(gi, k):
        h = outdeg_hist[gi]
        return sum(int(c) * comb(int(dv), k - 1) for dv, c in enumerate(h) if c and dv >= k - 1)
    srv = K.Server(workers=args.workers, max_batch=args.max_batch)
    gids = [None] * len(graphs)
    t = time.perf_counter()
    for i in range(args.resident):
        gids[i] = srv.add_graph(graphs[i], plans[i]["orient"])
    log(f"resident {args.resident} graphs in {time.perf_counter() - t:.1f}s", srv.stats())
    live = args.resident
    rates = [float(x) for x in (args.rates or str(args.rate)).split(",")]

    def stream(rate, nq, with_uploads):
        """Poisson (graph, k) queries at `rate` per second (0: all at once);
        the 64 uploads arrive during the first stream"""
        nonlocal live
        rng = np.random.default_rng(55)
        gaps = rng.exponential(1.0 / rate, nq) if rate > 0 else np.zeros(nq)
        pick = rng.random(nq)
        ks = rng.integers(3, 11, nq)
        modes = rng.integers(0, 2, nq)
        tickets, sent = [None] * nq, [None] * nq
        converted = 0

        def uploader():  # graphs arriving while queries run
            nonlocal live
            for i in range(args.resident, args.resident + args.uploads):
                time.sleep(max(gaps.sum(), 1.0) / (2 * args.uploads))
                gids[i] = srv.add_graph(graphs[i], plans[i]["orient"])
                live = i + 1

        srv.reset_stats()
        up = threading.Thread(target=uploader) if with_uploads else None
        t0 = time.perf_counter()
        if up:
            up.start()
        nxt = t0
        for i in range(nq):
            nxt += gaps[i]
            d = nxt - time.perf_counter()
            if d > 0:
                time.sleep(d)
            gi = int(pick[i] * live)
            algo = "pivot" if modes[i] else "orient"
            if algo == "orient" and orient_bound(gi, int(ks[i])) > ORIENT_BOUND:
                algo = "pivot"
                converted += 1
            q = dict(k=int(ks[i]), algo=algo, gran=plans[gi]["gran"], part_size=1)
            sent[i] = (gi, q)
            tickets[i] = srv.submit(gids[gi], **q)
            if (i + 1) % 200 == 0:
                log(f"rate {rate}: submitted {i + 1}/{nq}", srv.stats())
        if up:
            up.join()
        while srv.stats()["completed"] < nq:  # drain, printing progress
            time.sleep(1.0)
            log("draining", srv.stats())
        srv.drain()
        wall = time.perf_counter() - t0
        st = srv.stats()
        got = [srv.wait(tk).count for tk in tickets]
        log(f"rate {rate}", st)
        return sent, got, st, wall, converted

    sweep = []
    first = None
    for ri, rate in enumerate(rates):
        sent, got, st, wall, converted = stream(rate, args.queries, ri == 0)
        sweep.append({"offered_rate": rate if rate > 0 else "all at once", "queries_per_s": st["queries_per_s"],
                      "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
                      "mean_batch": st["mean_batch"], "resident_admitted": st["resident_admitted"],
                      "joined_running": st["joined_running"], "resident_launches": st["resident_launches"],
                      "wall_s": wall})
        if first is None:
            first = (sent, got, st, wall, converted)
    sent, got, st, wall, converted = first
    # per-call plain kernels on the first stream's queries, back to back
    ors = {}
    t = time.perf_counter()
    plain = []
    each = []
    nq = len(sent)
    for i, (gi, q) in enumerate(sent):
        if gi not in ors:
            ors[gi] = graphs[gi].orient(plans[gi]["orient"])
        t1 = time.perf_counter()
        plain.append(ors[gi].count(**q).count)
        each.append((time.perf_counter() - t1) * 1e3)
        if (i + 1) % 200 == 0:
            log(f"plain {i + 1}/{nq}")
    t_plain = time.perf_counter() - t
    same = got == plain
    log(f"plain back-to-back {t_plain:.2f}s; identical={same}")
    srv.close()
    if not same:
        raise SystemExit("served counts differ from the per-call counts")
    out = {"metric": "serving p50/p99 vs per-call plain kernels", "value": st["queries_per_s"],
           "unit": "queries/s", "n_gpus": 1, "higher_is_better": True, "dtype": "u64", "data": "synthetic",
           "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
           "config": {"workload": f"{args.resident} resident graphs + {args.uploads} uploaded mid-run (ER, Chung-Lu, "
                                  f"R-MAT, BA; <=2M edges); {nq} Poisson queries per rate, k~U{{3..10}}, "
                                  "orient|pivot, arrival seed 55",
                      "offered_rate": rates[0], "wall_s": wall, "batches": st["batches"],
                      "orient_queries_sent_as_pivot": converted, "orient_bound": ORIENT_BOUND,
                      "mean_batch": st["mean_batch"], "workers": args.workers,
                      "resident_admitted": st["resident_admitted"], "joined_running": st["joined_running"],
                      "resident_launches": st["resident_launches"], "sweep": sweep,
                      "plain_back_to_back_s": t_plain, "plain_queries_per_s": nq / t_plain,
                      "plain_ms_p50": float(np.percentile(each, 50)), "plain_ms_p99": float(np.percentile(each, 99)),
                      "plain_ms_max": float(max(each)),
                      "heaviest": [dict(kind=specs[sent[j][0]]["kind"], k=sent[j][1]["k"], algo=sent[j][1]["algo"],
                                        ms=round(each[j], 1)) for j in np.argsort(each)[-5:][::-1]],
                      "resident_bytes": st["resident_bytes"]}}
    print(json.dumps(out), flush=True)


def run_config(args):
    """The other single-graph configurations, counted per (orientation,
    granularity) with every result cross-checked:
      c1: Erdos-Renyi G(n=10,000, m=100,000), seed 1, degree orientation,
          node tasks, orientation DFS with P = 8 (-q o8b), k = 3 and 4;
      c3: R-MAT scale 22, edge factor 16, (a,b,c) = (.57,.19,.19), seed 3,
          {degree, degen} x {node, edge} pivoting, k = 3..11 (--ks)."""
    import numpy as np
    import kcgpu_amd as K
    if args.workload == "c1":
        G = K.Graph.generate(K.GEN_ER, seed=1, n=10000, m=100000)
        combos, ks, algo, P = [("degree", "node")], [3, 4], "orient", 8
    else:
        G = K.Graph.generate(K.GEN_RMAT, seed=3, scale=22, edge_factor=16, a=0.57, b=0.19, c=0.19)
        combos = [tuple(c.split(":")) for c in args.combos.split(",")]
        ks, algo, P = [int(x) for x in args.c3_ks.split(",")], "pivot", 1
    gi = G.info()
    log("graph", gi)
    ors, rows = {}, []
    for o, g in combos:
        if o not in ors:
            ors[o] = G.orient(o)
            log("oriented", o, ors[o].info())
        for k in ks:
            m_or = ors[o].info()["m"]
            if args.chunks > 1:
                # the count as the sum of its task residues mod `chunks`
                # (kc_count_sampled), one result line per residue, for counts
                # too long for one call; --chunk-range a:b takes residues a..b-1
                a0, b0 = (int(x) for x in (args.chunk_range or f"0:{args.chunks}").split(":"))
                tot, ovf, ms = 0, False, 0.0
                for off in range(a0, b0):
                    r = ors[o].count_sampled(k, args.chunks, off, algo, g, part_size=P, check=False)
                    tot += r.count
                    ovf |= r.overflow
                    ms += r.raw["count_ms"]
                    log(json.dumps({"orient": o, "gran": g, "k": k, "stride": args.chunks, "offset": off,
                                    "count": r.count, "overflow": r.overflow, "device_ms": round(r.raw["count_ms"], 3)}))
                ovf |= tot >= 2 ** 64
                row = {"orient": o, "gran": g, "k": k, "residues": f"{a0}:{b0} of {args.chunks}", "count": tot,
                       "overflow": ovf, "device_ms": round(ms, 3)}
                rows.append(row)
                log(json.dumps(row))
                continue
            r = ors[o].count(k, algo, g, part_size=P, check=False)
            ms = r.raw["count_ms"]
            row = {"orient": o, "gran": g, "k": k, "count": r.count, "overflow": r.overflow, "device_ms": round(ms, 3),
                   "kcliques_per_s": r.count / (ms / 1e3) if ms else None,
                   "edges_per_s": m_or / (ms / 1e3) if ms else None}
            rows.append(row)
            log(json.dumps(row))
    agree = all(len({(x["count"], x["overflow"]) for x in rows if x["k"] == k}) == 1 for k in ks)
    print(json.dumps({"workload": args.workload, "graph": gi, "agree_across_modes": agree,
                      "orient_ms": {o: round(v.info()["preprocess_ms"], 3) for o, v in ors.items()},
                      "max_core": {o: v.info()["max_core"] for o, v in ors.items()}, "counts": rows}), flush=True)
    if not agree:
        sys.exit(1)


def run_lkc(args):
    """-m lkc on the headline graph: per-vertex counts at each k (--ks),
    checked by sum_v local[v] = k * count against the plain count."""
    import numpy as np
    import kcgpu_amd as K
    G = K.Graph.generate(K.GEN_CHUNG_LU, seed=SEED, **CL)
    Og = G.orient("degen")
    rows = []
    for k in [int(x) for x in args.ks.split(",")]:
        loc, r = Og.count_local(k, "pivot", "edge", part_size=1, check=False)
        plain = Og.count(k, "pivot", "edge", part_size=1, check=False)
        ok = int(loc.sum(dtype=np.uint64)) == (k * plain.count) % (1 << 64) and r.count == plain.count
        row = {"k": k, "count": r.count, "device_ms": round(r.raw["count_ms"], 3),
               "plain_device_ms": round(plain.raw["count_ms"], 3), "sum_check": ok,
               "max_local": int(loc.max()), "vertices_in_a_clique": int((loc > 0).sum())}
        rows.append(row)
        log(json.dumps(row))
        if not ok:
            sys.exit(1)
    print(json.dumps({"workload": "lkc", "graph": "chung-lu n=1.7M m=11.1M seed 2, degen, edge pivoting",
                      "per_k": rows}), flush=True)


def run_formats(args):
    """On-disk formats on the headline graph (host only): write it as TSV,
    convert tsv-bel and bel-mtx, load each, and load the oriented-CSR cache."""
    import tempfile
    import numpy as np
    import kcgpu_amd as K
    G = K.Graph.generate(K.GEN_CHUNG_LU, seed=SEED, **CL)
    rp, col = G.csr()
    src = np.repeat(np.arange(len(rp) - 1, dtype=np.uint32), np.diff(rp).astype(np.int64))
    keep = src < col
    out = {}
    with tempfile.TemporaryDirectory() as d:
        tsv, bel, mtx, kco = (os.path.join(d, "g" + e) for e in (".tsv", ".bel", ".mtx", ".kco"))
        w = np.empty((int(keep.sum()), 3), dtype=np.uint64)
        w[:, 0], w[:, 1], w[:, 2] = col[keep], src[keep], 1
        import pandas as pd
        pd.DataFrame(w).to_csv(tsv, sep=" ", header=False, index=False)
        sums = {}
        for name, fn in (("tsv-bel", lambda: K.convert("tsv-bel", tsv, bel)),
                         ("bel-mtx", lambda: K.convert("bel-mtx", bel, mtx)),
                         ("load_tsv", lambda: K.Graph.load(tsv)), ("load_bel", lambda: K.Graph.load(bel)),
                         ("load_mtx", lambda: K.Graph.load(mtx))):
            t = time.perf_counter()
            r = fn()
            out[name + "_s"] = round(time.perf_counter() - t, 3)
            if r is not None:
                sums[name] = r.info()["checksum"]
        out["same_graph_from_every_format"] = len(set(sums.values()) | {G.info()["checksum"]}) == 1
        out["tsv_bytes"], out["bel_bytes"], out["mtx_bytes"] = (os.path.getsize(x) for x in (tsv, bel, mtx))
        t = time.perf_counter()
        Og = K.Graph.load(bel).orient("degen")
        out["load_bel_and_orient_s"] = round(time.perf_counter() - t, 3)
        Og.save(kco)
        t = time.perf_counter()
        L = K.Oriented.load(kco)
        out["load_kco_s"] = round(time.perf_counter() - t, 3)
        out["kco_bytes"] = os.path.getsize(kco)
        out["kco_count_k5_equal"] = L.count(5, "pivot", "edge").count == Og.count(5, "pivot", "edge").count
    print(json.dumps({"workload": "formats", "graph": "chung-lu n=1.7M m=11.1M seed 2", **out}), flush=True)


def run_copy(args):
    """Counter calibration: device-to-device copies of a known size through the
    HIP runtime's copy kernel, so rocprofv3's FETCH_SIZE / WRITE_SIZE can be
    scaled to bytes (see profiles/)."""
    import ctypes
    hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    nbytes = args.copy_mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    t = time.perf_counter()
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    dt = time.perf_counter() - t
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"workload": "copy", "bytes_per_copy": nbytes, "copies": 3,
                      "GBps_read_plus_write": 3 * 2 * nbytes / dt / 1e9}), flush=True)


def host_cores():
    """(threads the CPU leg uses, logical CPUs this process may run on,
    physical cores of the host from lscpu or None).  The CPU leg keeps to the
    16 host threads a one-GPU share of the machine gives."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    phys = None
    try:
        import subprocess
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        phys = len({l for l in out.splitlines() if l and not l.startswith("#")}) or None
    except Exception:
        pass
    return min(16, usable), usable, phys


def cpu_baseline(K, Og, ks, m_or, stride, repeats):
    """The oracle's enumerator (pivoting with the same closed-form tails, C
    with OpenMP) over every stride-th oriented edge task of the same oriented
    CSR, timed as the median of `repeats` runs; the GPU count over the same
    tasks (kc_count_sampled) must equal the oracle's at every k."""
    import numpy as np
    from oracle import oracle as O
    orp, oc, _ = Og.export()
    nt, usable, phys = host_cores()
    runs, counts = [], {}
    for rep in range(repeats):
        t = time.perf_counter()
        for k in ks:
            r = O.count(orp, oc, k, "pivot_tails", "edge", stride=stride, offset=0, threads=nt)
            counts[k] = r
        runs.append(time.perf_counter() - t)
        log(f"cpu run {rep}: {runs[-1]:.2f}s")
    tc = float(np.median(runs))
    parity = {}
    for k in ks:
        g = Og.count_sampled(k, stride, 0, "pivot", "edge")
        parity[str(k)] = {"gpu": g.count, "oracle": counts[k]["count"], "tasks": counts[k]["tasks"]}
        if g.count != counts[k]["count"] or g.raw["tasks"] != counts[k]["tasks"] or g.overflow:
            raise SystemExit(f"sampled count differs from the oracle at k={k}: gpu {g.count} "
                             f"({g.raw['tasks']} tasks) oracle {counts[k]['count']} ({counts[k]['tasks']} tasks)")
    tasks = sum(counts[k]["tasks"] for k in ks)
    sample_edges = (m_or + stride - 1) // stride
    return {"value": len(ks) * sample_edges / tc, "unit": "edges/s", "cores": nt, "kind": "port",
            "sample": f"every {stride}th oriented edge ({sample_edges} edges, {tasks} tasks with |T|>=k-2) "
                      f"at k={','.join(map(str, ks))}; median of {repeats} runs on {nt} OpenMP threads",
            "seconds": round(tc, 2), "runs_s": [round(x, 2) for x in runs],
            "host": {"usable_logical_cpus": usable, "physical_cores": phys},
            "parity": {"stride": stride, "offset": 0, "per_k": parity, "equal": True}}


def sum_residues(paths):
    """The full count from the per-residue lines of `--workload c3 --chunks N`
    runs (one JSON line per residue, possibly over several files and calls):
    one result per (orientation, granularity, k, N), with the residues that are
    missing or repeated with different counts named; the sum overflows when a
    residue did or the total reaches 2^64."""
    res = {}
    for path in paths:
        with open(path) as f:
            for line in f:
                if not line.startswith('{"orient"') or '"offset"' not in line:
                    continue
                d = json.loads(line)
                key = (d["orient"], d["gran"], d["k"], d["stride"])
                prev = res.setdefault(key, {}).get(d["offset"])
                cur = (d["count"], bool(d["overflow"]), d["device_ms"])
                if prev is not None and prev[:2] != cur[:2]:
                    raise ValueError(f"residue {d['offset']} of {key} has two counts: {prev[0]} and {cur[0]}")
                res[key][d["offset"]] = cur
    out = []
    for (o, g, k, n), r in sorted(res.items()):
        tot = sum(c for c, _, _ in r.values())
        out.append({"orient": o, "gran": g, "k": k, "residues_mod": n, "have": len(r),
                    "missing": sorted(set(range(n)) - set(r)), "count": tot,
                    "overflow": any(v for _, v, _ in r.values()) or tot >= 2 ** 64,
                    "device_s": round(sum(m for _, _, m in r.values()) / 1e3, 3)})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--ks", default="5,6,7,8")
    ap.add_argument("--part-size", type=int, default=0, help="0 = autotune")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=60)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--workload", choices=["c2", "batch", "serve", "copy", "c1", "c3", "lkc", "formats"], default="c2",
                    help="c2: the headline k=5..8 count (default); batch: the 4096-graph batched mix; "
                         "serve: the serving loop; c1 / c3: the ER and R-MAT configurations; "
                         "lkc: per-vertex counts; formats: converters and the oriented-CSR cache")
    ap.add_argument("--combos", default="degen:edge,degen:node,degree:edge,degree:node")
    ap.add_argument("--c3-ks", default="3,4,5,6,7,8,9,10,11")
    ap.add_argument("--chunks", type=int, default=1, help="c3: count as the sum over task residues mod N")
    ap.add_argument("--chunk-range", default="", help="c3 with --chunks: residues a:b only")
    ap.add_argument("--sum-residues", nargs="+", default=None, metavar="LOG",
                    help="sum the residue lines of --chunks runs saved in LOG files (no GPU)")
    ap.add_argument("--graphs", type=int, default=4096)
    ap.add_argument("--resident", type=int, default=512)
    ap.add_argument("--uploads", type=int, default=64)
    ap.add_argument("--queries", type=int, default=600)
    ap.add_argument("--rate", type=float, default=8.0)
    ap.add_argument("--rates", default="", help="serving: comma list of offered rates (0 = all at once)")
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--max-batch", type=int, default=256)
    args = ap.parse_args()
    if args.sum_residues:
        rows = sum_residues(args.sum_residues)
        for r in rows:
            print(json.dumps(r), flush=True)
        sys.exit(0 if rows and all(not r["missing"] for r in rows) else 1)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X only: there is no multi-GPU path (the tasks would shard by "
              "edge, each GPU holding the oriented CSR, with one 8-byte sum at the end). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    ks = [int(x) for x in args.ks.split(",")]
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload == "copy":
        return run_copy(args)
    if args.workload in ("c1", "c3"):
        return run_config(args)
    if args.workload == "lkc":
        return run_lkc(args)
    if args.workload == "formats":
        return run_formats(args)

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
    dom_cls = {k: {"ms": [[] for _ in range(6)], "bytes": [[] for _ in range(6)], "wave_ms": [[] for _ in range(6)]}
               for k in ks}
    t_all = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        last = step()
        times.append(time.perf_counter() - t)
        for k in ks:
            dev_ms[k].append(last[k].raw["count_ms"])
        for k in ks:
            for c in range(6):
                dom_cls[k]["ms"][c].append(last[k].raw["class_ms"][c])
                dom_cls[k]["bytes"][c].append(last[k].raw["class_bytes"][c])
                dom_cls[k]["wave_ms"][c].append(last[k].raw["class_task_ms_sum"][c])
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
                         "tree_nodes": last[k].raw["tree_nodes"], "part_size": last[k].raw["part_size"],
                         # the pivoting kernels run one task per 64-lane wave; the
                         # partition size shapes only the orientation DFS
                         "lanes_per_task": 64}

    # dominant kernel symbol: the size class whose tasks hold the device's
    # waves longest over the step (summed per-task wave time; the launches of
    # the small classes overlap, so their durations all span most of a
    # count), with its launch duration averaged over its launches
    cls_ms = np.zeros(6)
    cls_bytes = np.zeros(6)
    cls_wave = np.zeros(6)
    cls_launch = np.zeros(6)
    cls_var = [0] * 6
    for k in ks:
        for c in range(6):
            if last[k].raw["class_kernel"][c]:
                cls_ms[c] += np.mean([x for x in dom_cls[k]["ms"][c]])
                cls_bytes[c] += np.mean([x for x in dom_cls[k]["bytes"][c]])
                cls_wave[c] += np.mean([x for x in dom_cls[k]["wave_ms"][c]])
                cls_launch[c] += 1
                cls_var[c] = last[k].raw["class_kernel"][c]
    dc = int(np.argmax(cls_wave))
    avg_ms = cls_ms[dc] / max(cls_launch[dc], 1)
    avg_bytes = cls_bytes[dc] / max(cls_launch[dc], 1)
    achieved = avg_bytes / (avg_ms / 1e3) / 1e9 if avg_ms else 0.0
    # HBM traffic and issue rates of the dominant kernel come from rocprofv3
    # counter runs of this same command (profiles/): counters cannot be read
    # inside a timed run
    traffic, issue = None, None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        traffic = tj.get("bytes_per_launch")
        issue = {"valu_issue_frac": tj.get("valu_issue_frac"),
                 "valu_insts_per_simd_cycle": tj.get("valu_insts_per_simd_cycle"),
                 "valu_issue_ceiling": 0.25,
                 "salu_insts_per_cu_cycle": tj.get("salu_insts_per_cu_cycle"),
                 "lds_bank_conflict_frac": tj.get("lds_bank_conflict_frac"),
                 "lds_conflict_cycles_per_lds_inst": tj.get("lds_conflict_cycles_per_lds_inst"),
                 "kernel": tj.get("kernel"), "source": tj.get("source")}
    names = {1: "k_count<pivot,LDS>", 2: f"k_pivot_reg<{[1, 2, 4][dc] if dc < 3 else '?'}>", 3: "k_count<pivot,HBM>",
             4: "k_pivot_wide<LDS>", 5: "k_pivot_wide<HBM>"}
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": f"{names.get(cls_var[dc], '?')} (size class {dc}, tasks of <= {[64, 128, 256, 512, 1024, 'more'][dc]} vertices)",
                "launches_per_step": int(cls_launch[dc]), "avg_launch_ms": round(avg_ms, 3),
                "alg_bytes_per_launch": int(avg_bytes), "issue": issue,
                "note": "instruction-issue-bound (VALU + SALU) SCT on a dense core: algorithmic HBM bytes are tiny "
                        "next to the bitmap work; issue.valu_issue_frac is the bound that governs this kernel"}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(K, Og, ks, m_or, args.cpu_stride, args.cpu_repeats)
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
