"""Serving through the resident launch (continuous batching), against kc_count
alone and the oracle, through the C ABI."""
import time

import numpy as np
import pytest

import kcgpu_amd as K
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RMAT = dict(edge_factor=16, a=0.57, b=0.19, c=0.19)


def _wait_in_order(srv, tickets, limit_s=120.0):
    """poll until every ticket is done; (completion order, results by ticket)"""
    order, res = [], {}
    pending = set(tickets)
    t0 = time.perf_counter()
    t_note = t0
    while pending and time.perf_counter() - t0 < limit_s:
        for t in list(pending):
            r = srv.poll(t)
            if r is not None:
                order.append(t)
                res[t] = r
                pending.discard(t)
        if time.perf_counter() - t_note > 10.0:
            t_note = time.perf_counter()
            print(f"{t_note - t0:.0f} s: {len(pending)} of {len(tickets)} queries pending", flush=True)
        time.sleep(0.0005)
    assert not pending, f"{len(pending)} queries still pending after {limit_s} s"
    return order, res


def test_short_queries_join_running_work_and_finish_first():
    """A long query (R-MAT scale 16, degeneracy order, edge tasks at k=9;
    tasks above 64 vertices, so a worker counts it with the size-class
    kernels) is on the GPU when 4 medium queries (Barabasi-Albert n=100,000,
    100,000 node tasks: the bulk queue of the resident launch) and 24 short
    ones (n=3,000: the express queue) arrive.  The short and medium queries
    join the running resident launches and finish before the long query; every
    count equals kc_count on the same oriented graph alone, and the short
    ones equal the oracle."""
    longG = K.Graph.generate(K.GEN_RMAT, seed=3, scale=16, **RMAT)
    med = K.Graph.generate(K.GEN_BA, seed=99, n=100000, attach=20)
    shorts = [K.Graph.generate(K.GEN_BA, seed=100 + i, n=3000, attach=6 + i % 5) for i in range(6)]
    srv = K.Server(workers=2)
    try:
        gl = srv.add_graph(longG, "degen")
        gm = srv.add_graph(med, "degen")
        gs = [srv.add_graph(G, "degen") for G in shorts]
        tl = srv.submit(gl, 9, "pivot", "edge")
        time.sleep(0.05)  # the long query is on the GPU
        mq = [(gm, k, "pivot", "node") for k in (5, 6, 7, 8)]
        tm = [srv.submit(*q) for q in mq]
        time.sleep(0.01)  # the medium queries are in the resident launch
        sq = [(gs[i % 6], 4 + i % 4, "pivot", ("node", "edge")[i % 2]) for i in range(24)]
        ts = [srv.submit(*q) for q in sq]
        order, res = _wait_in_order(srv, [tl] + tm + ts)
        st = srv.stats()
        print("order of the long query:", order.index(tl), "of", len(order), st, flush=True)
        assert order[-1] == tl
        assert st["resident_admitted"] == 28 and st["joined_running"] >= 24
        Ol = longG.orient("degen")
        assert res[tl].count == Ol.count(9, "pivot", "edge").count
        Om = med.orient("degen")
        for t, (_, k, a, gr) in zip(tm, mq):
            assert res[t].count == Om.count(k, a, gr).count, k
        ors = [G.orient("degen") for G in shorts]
        for t, (g, k, a, gr) in zip(ts, sq):
            i = gs.index(g)
            assert res[t].count == ors[i].count(k, a, gr).count, (i, k, gr)
        for i in range(3):
            rp, col, _ = ors[i].export()
            for k in (4, 5, 6, 7):
                exp = O.count(rp.astype(np.uint64), col, k, "pivot_tails", "node")["count"]
                got = [res[t].count for t, q in zip(ts, sq) if q[0] == gs[i] and q[1] == k]
                assert all(x == exp for x in got), (i, k, got, exp)
    finally:
        srv.close()


def test_graph_uploaded_mid_launch_and_mixed_modes():
    """Queries on a graph added while the resident launch runs (counted by the
    workers until the launch restarts), orientation-mode queries (workers)
    and pivoting queries (resident launch) interleaved: every count equals
    kc_count alone; then, once the server is idle, the new graph's queries go
    to the resident launch too."""
    base = [K.Graph.generate(K.GEN_CHUNG_LU, seed=20 + i, n=40000, m=200000, gamma=2.1, max_degree=800.0)
            for i in range(3)]
    late = K.Graph.generate(K.GEN_ER, seed=31, n=20000, m=150000)
    srv = K.Server(workers=2)
    try:
        gids = [srv.add_graph(G, "degen") for G in base]
        ors = [G.orient("degen") for G in base]
        qs = [(j % 3, 4 + j % 5, ("pivot", "orient")[j % 3 == 2], ("node", "edge")[j % 2]) for j in range(30)]
        tickets = [srv.submit(gids[i], k, a, gr) for i, k, a, gr in qs]
        glate = srv.add_graph(late, "degen")
        olate = late.orient("degen")
        t_late = [srv.submit(glate, k, "pivot", "edge") for k in (3, 4, 5)]
        order, res = _wait_in_order(srv, tickets + t_late)
        for t, (i, k, a, gr) in zip(tickets, qs):
            assert res[t].count == ors[i].count(k, a, gr, part_size=1).count, (i, k, a, gr)
        for t, k in zip(t_late, (3, 4, 5)):
            assert res[t].count == olate.count(k, "pivot", "edge").count
        srv.drain()
        time.sleep(0.01)
        before = srv.stats()["resident_admitted"]
        t2 = [srv.submit(glate, k, "pivot", "node") for k in (3, 4, 5)]
        _, res2 = _wait_in_order(srv, t2)
        assert srv.stats()["resident_admitted"] == before + 3
        for t, k in zip(t2, (3, 4, 5)):
            assert res2[t].count == olate.count(k, "pivot", "node").count
        print(srv.stats(), flush=True)
    finally:
        srv.close()



def _near_clique_blocks(nblk, size, dens, seed):
    """nblk disjoint near-cliques of `size` vertices (edge density dens)"""
    rng = np.random.default_rng(seed)
    a, b = np.triu_indices(size, 1)
    es, ed = [], []
    for i in range(nblk):
        keep = rng.random(a.size) < dens
        es.append(a[keep] + i * size)
        ed.append(b[keep] + i * size)
    return K.Graph.from_edges(np.concatenate(es).astype(np.uint32), np.concatenate(ed).astype(np.uint32),
                              nblk * size)


def test_express_restarts_while_bulk_runs():
    """A long query in the bulk resident launch (2,000 near-cliques of 62
    vertices, 124,000 node tasks, k=10) runs while short queries arrive one
    at a time with gaps longer than the idle stop, so the express launch
    stops and restarts (its queue restarting at index 0) several times next
    to the running bulk launch.  Every count equals kc_count alone."""
    bulkG = _near_clique_blocks(2000, 62, 0.9, seed=4)
    shorts = [K.Graph.generate(K.GEN_BA, seed=300 + i, n=3000, attach=5 + i) for i in range(4)]
    srv = K.Server(workers=2)
    try:
        gb = srv.add_graph(bulkG, "degen")
        gs = [srv.add_graph(G, "degen") for G in shorts]
        ob = bulkG.orient("degen")
        os_ = [G.orient("degen") for G in shorts]
        assert ob.info()["max_out_degree"] <= 64 and ob.info()["n"] > 65536
        t0 = time.perf_counter()
        tb = srv.submit(gb, 10, "pivot", "node")
        launches0 = srv.stats()["resident_launches"]
        done_short = []
        for i in range(8):
            t = srv.submit(gs[i % 4], 4 + i % 3, "pivot", "node")
            _, r = _wait_in_order(srv, [t], limit_s=60.0)
            done_short.append((i, r[t], time.perf_counter() - t0))
            time.sleep(0.005)  # longer than the 2 ms idle stop
        bulk_running = srv.poll(tb) is None
        _, rb = _wait_in_order(srv, [tb], limit_s=100.0)
        st = srv.stats()
        print(f"bulk done after {time.perf_counter() - t0:.3f} s; short queries done at",
              [round(x[2], 3) for x in done_short], "bulk still running after them:", bulk_running, st, flush=True)
        assert st["resident_admitted"] == 9
        assert st["resident_launches"] - launches0 >= 2
        assert rb[tb].count == ob.count(10, "pivot", "node").count
        for i, r, _ in done_short:
            assert r.count == os_[i % 4].count(4 + i % 3, "pivot", "node").count, i
    finally:
        srv.close()


def test_remove_graph_while_its_query_runs():
    """A graph removed while its query is in the bulk resident launch: the
    query still completes with kc_count's count, the graph's memory is
    freed (resident bytes fall back) by a worker after that query, and other
    queries keep completing meanwhile; a graph removed with nothing in
    flight is freed at once."""
    bulkG = _near_clique_blocks(1100, 62, 0.9, seed=5)  # 68,200 node tasks: the bulk queue
    other = K.Graph.generate(K.GEN_BA, seed=400, n=3000, attach=6)
    idle = K.Graph.generate(K.GEN_BA, seed=401, n=3000, attach=7)
    srv = K.Server(workers=2)
    try:
        go = srv.add_graph(other, "degen")
        base = srv.stats()["resident_bytes"]
        gi = srv.add_graph(idle, "degen")
        srv.remove_graph(gi)  # nothing in flight: freed by the call
        assert srv.stats()["resident_bytes"] == base
        gb = srv.add_graph(bulkG, "degen")
        assert srv.stats()["resident_bytes"] > base
        ob, oo = bulkG.orient("degen"), other.orient("degen")
        tb = srv.submit(gb, 10, "pivot", "node")
        srv.remove_graph(gb)  # its query is in flight
        with pytest.raises(K.KCError) as ei:
            srv.submit(gb, 4)
        assert ei.value.code == K.KC_ERR_NOTFOUND
        shorts = []
        for i in range(6):
            t = srv.submit(go, 4 + i % 3, "pivot", "node")
            _, r = _wait_in_order(srv, [t], limit_s=60.0)
            shorts.append((i, r[t]))
            time.sleep(0.005)
        _, rb = _wait_in_order(srv, [tb], limit_s=100.0)
        assert rb[tb].count == ob.count(10, "pivot", "node").count
        for i, r in shorts:
            assert r.count == oo.count(4 + i % 3, "pivot", "node").count, i
        t0 = time.perf_counter()
        while srv.stats()["resident_bytes"] != base and time.perf_counter() - t0 < 30.0:
            time.sleep(0.01)
        st = srv.stats()
        print(f"removed graph freed {time.perf_counter() - t0:.3f} s after its query", st, flush=True)
        assert st["resident_bytes"] == base
        assert st["resident_admitted"] == 7
    finally:
        srv.close()
