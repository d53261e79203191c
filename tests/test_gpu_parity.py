"""GPU parity: the HIP path through the C ABI against the oracle and the golden counts.

All tests need an MI355X; they call libkcgpu.so (kcgpu_amd) and compare bit
for bit (integer counts, tolerance 0).
"""
from math import comb

import numpy as np
import pytest

import kcgpu_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRANS = ("node", "edge")
ALGOS = ("orient", "pivot")


def _graph(g):
    e = np.array(g["edges"], dtype=np.uint32).reshape(-1, 2)
    return K.Graph.from_edges(e[:, 0], e[:, 1], g["n"])


def _host_csr(G):
    rp, col = G.csr()
    return rp.astype(np.uint64), col


def test_device_is_gfx950():
    d = K.device_info()
    print(d, flush=True)
    assert d["arch"].startswith("gfx950")
    assert d["cu_count"] > 0


@pytest.mark.parametrize("mode", ["degree", "degen"])
def test_orientation_matches_oracle(golden, mode):
    for g in golden:
        G = _graph(g)
        rp, col = _host_csr(G)
        Og = G.orient(mode)
        orp, oc, key = Og.export()
        if mode == "degen":
            pr, mc = O.peel_priority(rp, col)
            assert (key == pr).all(), g["name"]
            assert Og.info()["max_core"] == mc == g["max_core"], g["name"]
            erp, ec = O.orient(rp, col, pr)
        else:
            assert (key == np.diff(rp)).all(), g["name"]
            erp, ec = O.orient(rp, col)
        assert (orp == erp).all() and (oc == ec).all(), g["name"]
        assert Og.info()["max_out_degree"] == int(np.diff(erp).max(initial=0))


def test_cooperative_peel_matches_oracle(golden, monkeypatch):
    """The single cooperative peel kernel (KC_PEEL_MODE=coop; the default is
    the stepped peel, covered by every other degen test) gives the same key."""
    monkeypatch.setenv("KC_PEEL_MODE", "coop")
    graphs = [_graph(g) for g in golden]
    graphs.append(K.Graph.generate(K.GEN_CHUNG_LU, seed=3, n=20000, m=150000, gamma=2.1, max_degree=2000))
    for G in graphs:
        rp, col = _host_csr(G)
        pr, mc = O.peel_priority(rp, col)
        Og = G.orient("degen")
        orp, oc, key = Og.export()
        assert (key == pr).all() and Og.info()["max_core"] == mc
        erp, ec = O.orient(rp, col, pr)
        assert (orp == erp).all() and (oc == ec).all()


@pytest.mark.parametrize("mode", ["degree", "degen"])
@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("gran", GRANS)
def test_counts_vs_golden(golden, mode, algo, gran):
    for g in golden:
        Og = _graph(g).orient(mode)
        for k in range(1, 11):
            if algo == "orient" and g["name"] == "K25" and k > 8:
                continue
            for P in (1, 8, 64):
                r = Og.count(k, algo, gran, part_size=P)
                assert r.count == g["counts"][k] and not r.overflow, (g["name"], mode, algo, gran, k, P)


def test_partition_sizes_all_agree(golden):
    g = next(x for x in golden if x["name"] == "planted_400")
    Og = _graph(g).orient("degen")
    for algo in ALGOS:
        for gran in GRANS:
            for P in (1, 2, 4, 8, 16, 32, 64):
                for k in (4, 7):
                    r = Og.count(k, algo, gran, part_size=P)
                    assert r.count == g["counts"][k], (algo, gran, P, k)


def _planted(n, m, sizes, seed):
    """sparse ER with planted cliques: out-degrees reach every size class"""
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


@pytest.mark.parametrize("mode", ["degree", "degen"])
def test_size_classes_vs_oracle(mode):
    # out-degrees above 64, 128, 256, 512 and 1024 exercise every LDS class and the HBM class
    G = _planted(6000, 30000, [80, 150, 300, 600, 1100], seed=5)
    rp, col = _host_csr(G)
    key = O.peel_priority(rp, col)[0] if mode == "degen" else None
    orp, oc = O.orient(rp, col, key)
    Og = G.orient(mode)
    assert Og.info()["max_out_degree"] > 1024
    for k, algos in ((3, ALGOS), (4, ALGOS), (7, ("pivot",))):
        exp = O.count(orp, oc, k, "pivot", "node")["count"]
        for algo in algos:
            for gran in GRANS:
                r = Og.count(k, algo, gran, part_size=8)
                print(mode, k, algo, gran, r.count, r.raw["count_ms"], flush=True)
                assert r.count == exp, (mode, k, algo, gran)


def test_register_classes_side_by_side_vs_oracle():
    # tasks in all three register classes (<= 64, <= 128, <= 256 vertices), whose
    # launches run concurrently on separate streams; repeated counts must agree
    # with the oracle every time
    G = _planted(4000, 16000, [100, 100, 200, 200], seed=11)
    rp, col = _host_csr(G)
    orp, oc = O.orient(rp, col, O.peel_priority(rp, col)[0])
    Og = G.orient("degen")
    for k in (5, 6):
        for gran in GRANS:
            exp = O.count(orp, oc, k, "pivot", gran)["count"]
            for rep in range(3):
                r = Og.count(k, "pivot", gran, part_size=1)
                print(k, gran, rep, r.count, r.raw["class_tasks"][:3], flush=True)
                assert all(t > 0 for t in r.raw["class_tasks"][:3]), r.raw["class_tasks"]
                assert r.count == exp, (k, gran, rep)


def _planted_near(n, m, specs, seed):
    """sparse ER with planted near-cliques (size, edge density): dense candidate
    sets of 65..256 vertices whose complements hold triangles and 4-cliques"""
    rng = np.random.default_rng(seed)
    G0 = K.Graph.generate(K.GEN_ER, seed=seed, n=n, m=m)
    rp, col = G0.csr()
    src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    es, ed = [src], [col]
    for s, dens in specs:
        vs = rng.choice(n, size=s, replace=False).astype(np.uint32)
        a, b = np.triu_indices(s, 1)
        keep = rng.random(a.size) < dens
        es.append(vs[a[keep]])
        ed.append(vs[b[keep]])
    return K.Graph.from_edges(np.concatenate(es), np.concatenate(ed), n)


def test_near_clique_closures_vs_oracle():
    # the h = k-4 complement closures (one-word and two-word rows), the
    # two-word compaction and the triangle walks all run on these subproblems
    G = _planted_near(3000, 12000, [(90, 0.92), (140, 0.85), (200, 0.9)], seed=7)
    rp, col = _host_csr(G)
    pr, _ = O.peel_priority(rp, col)
    orp, oc = O.orient(rp, col, pr)
    for k in (5, 6, 7):
        exp = O.count(orp, oc, k, "pivot_tails", "edge")["count"]
        for mode in ("degree", "degen"):
            Og = G.orient(mode)
            for gran in GRANS:
                r = Og.count(k, "pivot", gran, part_size=1)
                print(mode, k, gran, r.count, r.raw["count_ms"], flush=True)
                assert r.count == exp, (mode, k, gran, r.count, exp)


def test_wide_kernel_near_clique_closures_vs_oracle():
    # node tasks above 256 vertices go to the wide kernel; their h = k-4
    # entries of 129..256 candidates are compacted to four-word rows and
    # closed by the complement counts (complement density 0.2: triangles and
    # 4-cliques in the complement)
    G = _planted_near(3000, 12000, [(380, 0.8)], seed=9)
    rp, col = _host_csr(G)
    pr, _ = O.peel_priority(rp, col)
    orp, oc = O.orient(rp, col, pr)
    assert np.diff(orp).max() > 256
    for k in (5, 6):
        exp = O.count(orp, oc, k, "pivot_tails", "node")["count"]
        for mode in ("degree", "degen"):
            Og = G.orient(mode)
            for gran in GRANS:
                r = Og.count(k, "pivot", gran, part_size=1)
                print(mode, k, gran, r.count, r.raw["count_ms"], r.raw["class_kernel"], flush=True)
                assert r.count == exp, (mode, k, gran, r.count, exp)


@pytest.mark.parametrize("side_by_side", ["1", "0"])
@pytest.mark.parametrize("mode", ["degree", "degen"])
def test_near_clique_closed_forms_k5_to_11(near_cliques, mode, side_by_side, monkeypatch):
    """Complements of sparse graphs (triangles, K4, paths, stars, C4, paws,
    diamonds, random 4..7-vertex graphs): the exact counts follow from the
    complement's independence polynomial (tests/golden/make_golden.py).
    Tasks of up to 199 vertices (register classes), 299 (wide kernel, LDS) and
    1099 (wide kernel, HBM slices) at k = 5..11; counts of 2^64 or more must
    come back as overflow.  The register classes run side by side and in
    sequence."""
    from conftest import complement_edges
    monkeypatch.setenv("KC_SIDE_BY_SIDE", side_by_side)
    for g in near_cliques:
        a, b = complement_edges(g["n"], g["h_edges"])
        Og = K.Graph.from_edges(a, b, g["n"]).orient(mode)
        for k in range(5, 12):
            exp = g["counts"][k]
            for gran in GRANS:
                if gran == "edge" and g["n"] > 1000 and (k not in (7, 8) or mode == "degree" or side_by_side == "0"):
                    continue  # 350,000 edge tasks of 257..1099 vertices: 10 s per count; k=7 exact, k=8 overflow
                r = Og.count(k, "pivot", gran, part_size=1, check=False)
                print(g["name"], mode, side_by_side, k, gran, r.count, r.overflow, f"{r.raw['count_ms']:.1f} ms",
                      r.raw["class_tasks"], flush=True)
                if exp >= 2 ** 64:
                    assert r.overflow, (g["name"], mode, k, gran)
                else:
                    assert r.count == exp and not r.overflow, (g["name"], mode, k, gran, r.count, exp)


def test_measured_autotune_plan_counts():
    # a graph with out-degrees above 64: the plan is measured (orientation,
    # granularity, and at k <= 5 the algorithm and partition size as well);
    # counting with it gives the oracle's count
    G = K.Graph.generate(K.GEN_CHUNG_LU, seed=1, n=100000, m=600000, gamma=2.1, max_degree=5000.0)
    rp, col = _host_csr(G)
    orp, oc = O.orient(rp, col)
    for k in (4, 5, 6):
        plan = G.autotune(k)
        print(k, plan, flush=True)
        assert plan["orient"] in ("degree", "degen") and plan["gran"] in GRANS
        assert plan["part_size"] in (1, 8, 64) and plan["algo"] in ALGOS
        if k > 5:
            assert plan["algo"] == "pivot" and plan["part_size"] == 1
        r = G.orient(plan["orient"]).count(k, plan["algo"], plan["gran"], part_size=plan["part_size"])
        assert r.count == O.count(orp, oc, k, "pivot_tails", "edge")["count"], (k, plan)


@pytest.mark.parametrize("kind,kw", [
    (K.GEN_ER, dict(n=10000, m=100000)),  # config 1 shape
    (K.GEN_BA, dict(n=20000, attach=12)),
    (K.GEN_CHUNG_LU, dict(n=30000, m=200000, gamma=2.1, max_degree=1500)),
    (K.GEN_RMAT, dict(scale=13, edge_factor=16, a=0.57, b=0.19, c=0.19)),
])
def test_generated_graphs_vs_oracle(kind, kw):
    G = K.Graph.generate(kind, seed=11, **kw)
    rp, col = _host_csr(G)
    pr, _ = O.peel_priority(rp, col)
    for mode, key in (("degree", None), ("degen", pr)):
        orp, oc = O.orient(rp, col, key)
        Og = G.orient(mode)
        xr, xc, _ = Og.export()
        assert (xr == orp).all() and (xc == oc).all()
        for k in (3, 4, 5, 7):
            exp = O.count(orp, oc, k, "orient", "edge")
            for algo in ALGOS:
                for gran in GRANS:
                    r = Og.count(k, algo, gran, part_size=16)
                    assert r.count == exp["count"], (kind, mode, k, algo, gran)
                    if algo == "orient" and gran == "edge":
                        assert r.raw["tasks"] == exp["tasks"]


def test_overflow_flag_and_exact_boundary():
    n = 70
    a, b = np.triu_indices(n, 1)
    Og = K.Graph.from_edges(a.astype(np.uint32), b.astype(np.uint32), n).orient("degree")
    for gran in GRANS:
        r = Og.count(30, "pivot", gran, check=False)
        assert r.overflow
        with pytest.raises(K.KCError) as ei:
            Og.count(30, "pivot", gran)
        assert ei.value.code == K.KC_ERR_OVERFLOW
        r = Og.count(32, "pivot", gran, check=False)
        assert r.overflow
        r = Og.count(12, "pivot", gran)
        assert r.count == comb(70, 12) and not r.overflow
    n = 66  # C(66, 32) = 7.0e18 < 2^64: exact
    a, b = np.triu_indices(n, 1)
    Og = K.Graph.from_edges(a.astype(np.uint32), b.astype(np.uint32), n).orient("degen")
    for gran in GRANS:
        r = Og.count(32, "pivot", gran)
        assert r.count == comb(66, 32) and not r.overflow


def test_bad_queries_are_errors():
    Og = K.Graph.generate(K.GEN_ER, seed=1, n=100, m=300).orient("degree")
    for bad in (dict(k=0), dict(k=33), dict(k=4, part_size=3), dict(k=4, part_size=128)):
        kw = dict(algo="pivot", gran="node")
        kw.update(bad)
        with pytest.raises(K.KCError) as ei:
            Og.count(**kw)
        assert ei.value.code == K.KC_ERR_ARG


def test_empty_and_tiny_graphs():
    for n, src, dst in ((1, [], []), (2, [0], [1]), (3, [0, 1], [1, 2])):
        G = K.Graph.from_edges(np.array(src, np.uint32), np.array(dst, np.uint32), n)
        for mode in ("degree", "degen"):
            Og = G.orient(mode)
            for k in range(1, 5):
                exp = {1: n, 2: len(src)}.get(k, 0)
                for algo in ALGOS:
                    for gran in GRANS:
                        assert Og.count(k, algo, gran).count == exp


def _small_mix(rng, count):
    """uneven small graphs (BA and ER), as in the batched workload"""
    out = []
    for i in range(count):
        n = int(rng.integers(50, 3000))
        if i % 2 == 0:
            G = K.Graph.generate(K.GEN_BA, seed=1000 + i, n=n, attach=int(rng.integers(2, 12)))
        else:
            m = min(n * int(rng.integers(2, 20)) // 2, n * (n - 1) // 4)
            G = K.Graph.generate(K.GEN_ER, seed=1000 + i, n=n, m=m)
        out.append(G)
    return out


def test_batch_equals_alone():
    rng = np.random.default_rng(7)
    graphs = _small_mix(rng, 40)
    graphs.append(_planted(3000, 12000, [90, 200, 400], seed=3))  # bins above 64 vertices
    ors = [G.orient("degen" if i % 3 else "degree") for i, G in enumerate(graphs)]
    pairs = []
    for i, Og in enumerate(ors):
        for _ in range(2):  # the same graph twice, with different queries
            q = dict(k=int(rng.integers(1, 9)), algo=ALGOS[int(rng.integers(0, 2))],
                     gran=GRANS[int(rng.integers(0, 2))], part_size=int(2 ** rng.integers(0, 7)))
            pairs.append((Og, q))
    pairs.append((ors[-1], dict(k=6, algo="pivot", gran="edge", auto_tune=True)))
    alone = [Og.count(**q).count for Og, q in pairs]
    mixed, st = K.batch_count(pairs)
    print("batch", st, flush=True)
    assert [r.count for r in mixed] == alone
    # a second batch on the same scratch, smaller, still exact
    mixed2, _ = K.batch_count(pairs[:7])
    assert [r.count for r in mixed2] == alone[:7]


def test_batch_overflow_is_per_query():
    n = 70
    a, b = np.triu_indices(n, 1)
    K70 = K.Graph.from_edges(a.astype(np.uint32), b.astype(np.uint32), n).orient("degree")
    small = K.Graph.generate(K.GEN_ER, seed=3, n=500, m=3000).orient("degen")
    pairs = [(K70, dict(k=30, algo="pivot", gran="node")), (small, dict(k=3, algo="pivot", gran="edge")),
             (K70, dict(k=10, algo="pivot", gran="edge"))]
    res, _ = K.batch_count(pairs, check=False)
    assert res[0].overflow and not res[1].overflow and not res[2].overflow
    assert res[1].count == small.count(3, "orient", "edge").count
    assert res[2].count == comb(70, 10)


def test_serving_results_equal_alone():
    import threading
    rng = np.random.default_rng(11)
    graphs = _small_mix(rng, 12) + [_planted(2500, 9000, [70, 150], seed=4)]
    srv = K.Server(workers=2, max_batch=16)
    try:
        gids = [srv.add_graph(G, "degen" if i % 2 else "degree") for i, G in enumerate(graphs)]
        ors = {gid: G.orient("degen" if i % 2 else "degree") for i, (gid, G) in enumerate(zip(gids, graphs))}
        qs = []
        for _ in range(120):
            gi = int(rng.integers(0, len(gids)))
            algo = ALGOS[int(rng.integers(0, 2))]
            if gi == len(gids) - 1:
                algo = "pivot"  # enumerating C(150, 9) cliques one by one takes hours
            qs.append((gids[gi], int(rng.integers(1, 10)), algo, GRANS[int(rng.integers(0, 2))]))
        tickets = [None] * len(qs)

        def submit(lo, hi):  # several client threads, queries arrive while others run
            for i in range(lo, hi):
                g, k, a, gr = qs[i]
                tickets[i] = srv.submit(g, k, a, gr)

        th = [threading.Thread(target=submit, args=(i * 30, (i + 1) * 30)) for i in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        mixed = [srv.wait(t).count for t in tickets]
        alone = [ors[g].count(k, a, gr).count for g, k, a, gr in qs]
        assert mixed == alone
        st = srv.stats()
        print("serving", st, flush=True)
        assert st["completed"] == len(qs) and st["batches"] >= 1
        # a graph larger than the budget left is refused
        big = K.Graph.generate(K.GEN_ER, seed=5, n=200000, m=2000000)
        tiny = K.Server(budget_bytes=8 << 20, workers=1)
        try:
            with pytest.raises(K.KCError) as ei:
                tiny.add_graph(big)
            assert ei.value.code == K.KC_ERR_BUDGET
        finally:
            tiny.close()
        srv.remove_graph(gids[0])
        with pytest.raises(K.KCError) as ei:
            srv.submit(gids[0], 4)
        assert ei.value.code == K.KC_ERR_NOTFOUND
    finally:
        srv.close()


def test_cli_end_to_end(tmp_path, golden):
    """kcgpu -g file -m kc -o ... -k ... -p ... -q ...: the reference's flags
    and output lines, with the exact count."""
    import os
    import re
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kcgpu_amd", "bin", "kcgpu")
    g = next(x for x in golden if x["name"] == "planted_200")
    p = tmp_path / "g.mtx"
    lines = ["%%MatrixMarket matrix coordinate pattern symmetric", f"{g['n']} {g['n']} {len(g['edges'])}"]
    lines += [f"{b + 1} {a + 1}" for a, b in g["edges"]]  # lower triangle, 1-based
    p.write_text("\n".join(lines) + "\n")
    for orient, by, spec, k in (("degree", "node", "o8b", 5), ("degen", "edge", "p64b", 7), ("degen", "node", "p1n", 9)):
        r = subprocess.run([cli, "-g", str(p), "-d", "0", "-m", "kc", "-o", orient, "-k", str(k), "-p", by,
                            "-q", spec, "-w", "0"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        m = re.search(r"Counter = (\d+)", r.stdout)
        assert m and int(m.group(1)) == g["counts"][k], r.stdout
        assert re.search(r"MOHA %d k-clique" % g["counts"][k], r.stdout)
        assert f"kcliques={g['counts'][k]} k={k} overflow=0" in r.stdout
        if orient == "degen":
            assert f"Max Core = {g['max_core']}" in r.stdout


def test_cli_kcore(tmp_path, golden):
    """-m kcore (src/main.cu:248-261): the device peel alone; prints the
    reference's "Max Core" line (networkx core number) and counts nothing,
    whatever -o says."""
    import os
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kcgpu_amd", "bin", "kcgpu")
    for name in ("planted_400", "ba_500_10", "K25", "cycle_50"):
        g = next(x for x in golden if x["name"] == name)
        p = tmp_path / f"{name}.txt"
        p.write_text("".join(f"{b} {a}\n" for a, b in g["edges"]))
        for orient in ("degen", "degree"):
            r = subprocess.run([cli, "-g", str(p), "-m", "kcore", "-o", orient], capture_output=True, text=True,
                               timeout=120)
            assert r.returncode == 0, r.stderr
            assert f"Max Core = {g['max_core']}\n" in r.stdout, (name, r.stdout)
            assert "Preprocess time" in r.stdout and "kcliques=" not in r.stdout and "Counter" not in r.stdout


def test_cli_graph_stats(tmp_path):
    """-x: MatrixStats and the 10 x 10 structure grid of the oriented CSR,
    checked on a hand-worked graph: edge 5-6 and triangle 7-8-9 on ten
    vertices, degree order -> rows 5:[6] 7:[8,9] 8:[9]; row lengths
    {0: 7, 1: 2, 2: 1}, mean 0.4, sd sqrt(4.4/10) = 0.66; every entry is a
    quarter of the four (250 thousandths)."""
    import os
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kcgpu_amd", "bin", "kcgpu")
    p = tmp_path / "g.tsv"
    p.write_text("5 6\n7 8\n7 9\n8 9\n")
    r = subprocess.run([cli, "-g", str(p), "-m", "kc", "-o", "degree", "-k", "3", "-p", "node", "-q", "p1b",
                        "-x", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    i = out.index("4,10,10,1,0.4,0.7,1,2,2,2,1")
    grid = out[i + 1:i + 11]
    empty = ",".join("-" * 10) + ","
    want = [empty] * 10
    want[5] = "-,-,-,-,-,-,250,-,-,-,"
    want[7] = "-,-,-,-,-,-,-,-,250,250,"
    want[8] = "-,-,-,-,-,-,-,-,-,250,"
    assert grid == want, "\n".join(grid)
    assert "kcliques=1 k=3 overflow=0" in r.stdout


def test_oriented_cache_round_trip(tmp_path, golden):
    """save -> load gives the same oriented CSR, key and counts; the CLI counts
    a .kco file without orienting again."""
    import os
    import re
    import subprocess
    g = next(x for x in golden if x["name"] == "planted_200")
    e = np.array(g["edges"], dtype=np.uint32).reshape(-1, 2)
    G = K.Graph.from_edges(e[:, 0], e[:, 1], g["n"])
    for mode in ("degree", "degen"):
        Og = G.orient(mode)
        p = tmp_path / f"g_{mode}.kco"
        Og.save(p)
        L = K.Oriented.load(p)
        a, b = Og.export(), L.export()
        assert all(np.array_equal(x, y) for x, y in zip(a, b))
        assert L.info()["max_out_degree"] == Og.info()["max_out_degree"]
        assert L.info()["max_core"] == Og.info()["max_core"]
        for k in (4, 7):
            for gran in GRANS:
                assert L.count(k, "pivot", gran, part_size=8).count == g["counts"][k]
                assert L.count(k, "orient", gran, part_size=8).count == g["counts"][k]
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kcgpu_amd", "bin", "kcgpu")
    r = subprocess.run([cli, "-g", str(tmp_path / "g_degen.kco"), "-m", "kc", "-k", "6", "-p", "edge", "-q", "p8b"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"kcliques={g['counts'][6]} k=6 overflow=0" in r.stdout


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("gran", GRANS)
def test_local_counts_vs_golden(golden, algo, gran):
    """-m lkc: per-vertex counts equal networkx's for every vertex."""
    for g in golden:
        if algo == "orient" and g["name"] == "K25":
            continue
        G = _graph(g)
        for mode in ("degree", "degen"):
            Og = G.orient(mode)
            for k in (1, 2, 3, 4, 5, 6):
                for P in (1, 8):
                    loc, r = Og.count_local(k, algo, gran, part_size=P)
                    if k <= 2:
                        exp = [1] * g["n"] if k == 1 else np.bincount(
                            np.array(g["edges"], dtype=np.int64).ravel(), minlength=g["n"]).tolist()
                    else:
                        exp = g["local"][str(k)]
                    assert loc.tolist() == exp, (g["name"], mode, k, P)
                    assert r.count == g["counts"][k]


@pytest.mark.parametrize("mode", ["degree", "degen"])
def test_local_counts_size_classes_vs_oracle(mode):
    G = _planted(6000, 30000, [80, 150, 300, 600, 1100], seed=5)
    rp, col = _host_csr(G)
    key = O.peel_priority(rp, col)[0] if mode == "degen" else None
    orp, oc = O.orient(rp, col, key)
    Og = G.orient(mode)
    # (orientation credits every clique member one by one: C(1100, 4) adds at
    # k=4 would take minutes, so it runs at k=3 here)
    for k, algos in ((3, ALGOS), (4, ("pivot",)), (7, ("pivot",))):
        exp, ovf = O.count_local(orp, oc, k, "pivot_tails", "node")
        for algo in algos:
            for gran in GRANS:
                loc, r = Og.count_local(k, algo, gran, part_size=8)
                print(mode, k, algo, gran, r.raw["count_ms"], flush=True)
                assert (loc == exp).all() and not r.overflow, (mode, k, algo, gran)
    # a power-law graph: every vertex against the oracle, sum = k * count
    G = K.Graph.generate(K.GEN_CHUNG_LU, seed=11, n=30000, m=200000, gamma=2.1, max_degree=1500)
    rp, col = _host_csr(G)
    orp, oc = O.orient(rp, col, O.peel_priority(rp, col)[0])
    Og = G.orient("degen")
    for k in (5, 8):
        exp, _ = O.count_local(orp, oc, k, "pivot_tails", "edge")
        loc, r = Og.count_local(k, "pivot", "edge", part_size=1)
        assert (loc == exp).all()
        assert int(loc.sum(dtype=np.uint64)) == k * r.count == k * Og.count(k, "pivot", "edge").count


def test_local_counts_cli(tmp_path, golden):
    import os
    import re
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kcgpu_amd", "bin", "kcgpu")
    g = next(x for x in golden if x["name"] == "planted_200")
    p = tmp_path / "g.txt"
    p.write_text("".join(f"{b} {a}\n" for a, b in g["edges"]))  # "dst src", 0-based
    r = subprocess.run([cli, "-g", str(p), "-m", "lkc", "-o", "degen", "-k", "5", "-p", "node", "-q", "p8b"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Local clique counter for the first 200 Nodes:" in r.stdout
    rows = re.findall(r"^(\d+)\t(\d+)$", r.stdout, flags=re.M)
    assert [int(c) for _, c in rows] == g["local"]["5"][:50]
    assert f"kcliques={g['counts'][5]} k=5 overflow=0" in r.stdout


def test_orient_batch_equals_one_by_one(golden):
    """kc_orient_batch: degree graphs through one captured hipGraph over the
    concatenation, degen and empty graphs one by one; every handle equals
    what kc_orient gives, and counts on it are unchanged."""
    rng = np.random.default_rng(21)
    graphs = [_graph(g) for g in golden] + _small_mix(rng, 30)
    graphs.append(K.Graph.from_edges(np.zeros(0, np.uint32), np.zeros(0, np.uint32), 5))  # no edges
    graphs.append(K.Graph.from_edges(np.zeros(0, np.uint32), np.zeros(0, np.uint32), 0))  # no vertices
    modes = ["degen" if i % 3 == 0 else "degree" for i in range(len(graphs))]
    batch = K.orient_batch(graphs, modes)
    for i, (G, mode, B) in enumerate(zip(graphs, modes, batch)):
        A = G.orient(mode)
        for x, y in zip(A.export(), B.export()):
            assert np.array_equal(x, y), (i, mode)
        ia, ib = A.info(), B.info()
        assert (ia["n"], ia["m"], ia["max_out_degree"], ia["max_core"]) == (ib["n"], ib["m"], ib["max_out_degree"],
                                                                            ib["max_core"]), (i, mode)
        if ia["m"]:
            assert B.count(5, "pivot", "edge").count == A.count(5, "pivot", "edge").count
    # handles of one shared allocation stay valid when others are freed
    keep = batch[1]
    exp = graphs[1].orient(modes[1]).count(4, "orient", "node").count
    del batch
    assert keep.count(4, "orient", "node").count == exp
