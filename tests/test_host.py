"""CPU tests of the host half of the C ABI: loaders, CSR builder, generators,
-q grammar, error paths, and that libkcgpu.so exports every declared symbol.
No GPU call is made here (device entry points are only checked to fail
cleanly when no device is present).
"""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import kcgpu_amd as K
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---- ABI surface ---------------------------------------------------------------
def _declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "kc.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(kc_[a-z_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    syms = _declared_symbols()
    assert "kc_count" in syms and "kc_orient" in syms and "kc_load" in syms
    out = subprocess.run(["nm", "-D", "--defined-only", K.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (kc_[a-z_]+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    # and the Python mirror binds each of them
    assert sorted(K.SIGNATURES) == syms


def test_noalloc_scratch_size_matches_header():
    hdr = open(os.path.join(ROOT, "include", "kc.h")).read()
    m = re.search(r"#define KC_NOALLOC_SCRATCH_MIB (\d+)", hdr)
    assert m and int(m.group(1)) == K.NOALLOC_SCRATCH_MIB


# ---- loaders -----------------------------------------------------------------
def _edges(G):
    rp, col = G.csr()
    src = np.repeat(np.arange(rp.size - 1), np.diff(rp).astype(np.int64))
    return {(int(a), int(b)) for a, b in zip(src, col) if a < b}, rp.size - 1


def test_mtx_general_pattern(tmp_path):
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n% a comment\n5 5 6\n"
                 "1 2\n2 1\n2 3\n3 3\n4 1\n4 2\n")
    e, n = _edges(K.Graph.load(p))
    assert n == 5  # ids stay 1-based: vertex 0 is isolated
    assert e == {(1, 2), (2, 3), (1, 4), (2, 4)}


def test_mtx_symmetric_with_values(tmp_path):
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n4 4 4\n2 1 0.5\n3 1 1.5\n3 2 2\n4 3 1e3\n")
    e, n = _edges(K.Graph.load(p))
    assert e == {(1, 2), (1, 3), (2, 3), (3, 4)} and n == 5


def test_tsv_comments_blank_lines_and_both_directions(tmp_path):
    p = tmp_path / "g.tsv"
    p.write_text("# snap header\n% other comment\n0\t1\n1\t0\n\n1 2\n2\t2\n3 0\n")
    e, n = _edges(K.Graph.load(p))
    assert e == {(0, 1), (1, 2), (0, 3)} and n == 4
    q = tmp_path / "g.txt"
    q.write_text("5 6\n")
    e, n = _edges(K.Graph.load(q))
    assert e == {(5, 6)} and n == 7


def _bel(path, pairs):
    with open(path, "wb") as f:
        for s, d in pairs:
            f.write(struct.pack("<QQQ", d, s, 0))  # dst @0, src @8, weight @16


def test_bel_records(tmp_path):
    p = tmp_path / "g.bel"
    _bel(p, [(0, 1), (1, 0), (1, 2), (2, 2), (3, 1)])
    e, n = _edges(K.Graph.load(p))
    assert e == {(0, 1), (1, 2), (1, 3)} and n == 4


def test_loader_errors(tmp_path):
    with pytest.raises(K.KCError) as ei:
        K.Graph.load(tmp_path / "missing.mtx")
    assert ei.value.code == K.KC_ERR_IO
    bad = tmp_path / "g.csv"
    bad.write_text("0,1\n")
    with pytest.raises(K.KCError) as ei:
        K.Graph.load(bad)
    assert ei.value.code == K.KC_ERR_FORMAT
    m = tmp_path / "g.mtx"
    m.write_text("%%MatrixMarket matrix coordinate pattern general\n3 4 1\n1 2\n")
    with pytest.raises(K.KCError) as ei:
        K.Graph.load(m)
    assert ei.value.code == K.KC_ERR_FORMAT  # not square
    t = tmp_path / "g.tsv"
    t.write_text("1 2\nx y\n")
    with pytest.raises(K.KCError) as ei:
        K.Graph.load(t)
    assert ei.value.code == K.KC_ERR_FORMAT
    b = tmp_path / "g.bel"
    b.write_bytes(b"\0" * 25)
    with pytest.raises(K.KCError) as ei:
        K.Graph.load(b)
    assert ei.value.code == K.KC_ERR_FORMAT


# ---- the reference's raw read loop (documentation of its behaviour) -----------
def test_reference_reader_symmetric_mtx_keeps_one_direction(tmp_path):
    # include/FIleReader.cuh:218-231: symmetric entries are stored as (col,row)
    # and the slot after each one is left as it was ((0,0) in a fresh batch)
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n4 4 3\n2 1\n3 1\n4 3\n")
    s, d = K.read_edges_reference(p)
    assert list(zip(s.tolist(), d.tolist())) == [(1, 2), (0, 0), (1, 3), (0, 0), (3, 4), (0, 0)]


def test_reference_reader_tsv_and_bel(tmp_path):
    p = tmp_path / "g.tsv"
    p.write_text("% c\n0 1\n2 2\n1 2\n")  # "dst src" per line, both directions emitted
    s, d = K.read_edges_reference(p)
    assert list(zip(s.tolist(), d.tolist())) == [(1, 0), (0, 1), (2, 1), (1, 2)]
    b = tmp_path / "g.bel"
    _bel(b, [(0, 1), (2, 1)])  # BEL is not symmetrized
    s, d = K.read_edges_reference(b)
    assert list(zip(s.tolist(), d.tolist())) == [(0, 1), (2, 1)]


def test_reference_reader_general_mtx_self_loop_leaves_slot(tmp_path):
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n3 3 3\n1 2\n2 2\n2 3\n")
    s, d = K.read_edges_reference(p)
    assert list(zip(s.tolist(), d.tolist())) == [(1, 2), (0, 0), (2, 3)]


def test_reference_reader_refuses_tsv_blank_line(tmp_path):
    p = tmp_path / "g.tsv"
    p.write_text("0 1\n\n1 2\n")  # the reference re-reads a stale line forever here
    with pytest.raises(K.KCError) as ei:
        K.read_edges_reference(p)
    assert ei.value.code == K.KC_ERR_FORMAT


# ---- CSR builder and generators -------------------------------------------------
def test_builder_matches_oracle_on_random_edges():
    rng = np.random.default_rng(3)
    for _ in range(5):
        n = int(rng.integers(2, 3000))
        m = int(rng.integers(1, 20000))
        s = rng.integers(0, n, m).astype(np.uint32)
        d = rng.integers(0, n, m).astype(np.uint32)
        G = K.Graph.from_edges(s, d, n)
        rp, col = G.csr()
        erp, ecol = O.csr_from_edges(s, d, n)
        assert (rp == erp).all() and (col == ecol).all()
        G2 = K.Graph.from_csr(rp, col)
        assert G2.info()["checksum"] == G.info()["checksum"]


def test_generators_are_deterministic_and_sized():
    cases = [(K.GEN_ER, dict(n=2000, m=15000)), (K.GEN_BA, dict(n=3000, attach=7)),
             (K.GEN_CHUNG_LU, dict(n=20000, m=90000, gamma=2.1, max_degree=800.0)),
             (K.GEN_RMAT, dict(scale=11, edge_factor=8, a=0.57, b=0.19, c=0.19))]
    for kind, kw in cases:
        a = K.Graph.generate(kind, seed=9, **kw).info()
        b = K.Graph.generate(kind, seed=9, **kw).info()
        c = K.Graph.generate(kind, seed=10, **kw).info()
        assert a == b and a["checksum"] != c["checksum"]
    assert K.Graph.generate(K.GEN_ER, seed=1, n=10000, m=100000).info()["m"] == 100000
    assert K.Graph.generate(K.GEN_CHUNG_LU, seed=1, n=20000, m=90000, gamma=2.1,
                            max_degree=800.0).info()["m"] == 90000
    ba = K.Graph.generate(K.GEN_BA, seed=1, n=3000, attach=7).info()
    assert ba["m"] == 7 + (3000 - 8) * 7


def test_generator_checksums_pinned():
    """The benchmark inputs are defined by these generators; pin a small
    instance of each so an unintended change of the byte stream shows."""
    got = {
        "er": K.Graph.generate(K.GEN_ER, seed=1, n=10000, m=100000).info()["checksum"],
        "ba": K.Graph.generate(K.GEN_BA, seed=42, n=3000, attach=8).info()["checksum"],
        "rmat": K.Graph.generate(K.GEN_RMAT, seed=3, scale=12, edge_factor=8, a=.57, b=.19, c=.19).info()["checksum"],
    }
    assert got == PINNED, got


PINNED = {"er": 8325718519874903858, "ba": 16771397143456052198, "rmat": 12736577283908619219}


def test_bad_generator_parameters():
    for kw in (dict(kind=K.GEN_ER, n=10, m=100), dict(kind=K.GEN_BA, n=5, attach=5),
               dict(kind=K.GEN_RMAT, scale=0, edge_factor=4), dict(kind=9)):
        kind = kw.pop("kind")
        with pytest.raises(K.KCError) as ei:
            K.Graph.generate(kind, seed=1, **kw)
        assert ei.value.code == K.KC_ERR_ARG


# ---- -q grammar (include/Config.h:235-271) ---------------------------------------
def test_spec_grammar():
    assert K.parse_spec("o8b") == {"algo": "orient", "part_size": 8, "binary": True}
    assert K.parse_spec("p64n") == {"algo": "pivot", "part_size": 64, "binary": False}
    assert K.parse_spec("p8") == {"algo": "pivot", "part_size": 8, "binary": True}  # the reference ignores this one
    assert K.parse_spec("o1h")["part_size"] == 1
    for bad in ("", "o", "x8b", "o3b", "o128b", "o8q", "p8bb", "pb"):
        with pytest.raises(K.KCError) as ei:
            K.parse_spec(bad)
        assert ei.value.code == K.KC_ERR_ARG, bad


def test_device_calls_fail_cleanly_without_gpu():
    try:
        K.device_info()
        pytest.skip("a GPU is present")
    except K.KCError as e:
        assert e.code == K.KC_ERR_DEVICE
    G = K.Graph.generate(K.GEN_ER, seed=1, n=100, m=200)
    with pytest.raises(K.KCError) as ei:
        G.orient("degen")
    assert ei.value.code == K.KC_ERR_DEVICE
    with pytest.raises(K.KCError) as ei:
        K.Server()
    assert ei.value.code == K.KC_ERR_DEVICE
    # the device CSR builder and the edge-list orientation need the device too
    s = np.array([0, 1, 2], np.uint32)
    d = np.array([1, 2, 0], np.uint32)
    with pytest.raises(K.KCError) as ei:
        K.Graph.from_edges_device(s, d)
    assert ei.value.code == K.KC_ERR_DEVICE
    with pytest.raises(K.KCError) as ei:
        K.orient_edges(s, d, "degen")
    assert ei.value.code == K.KC_ERR_DEVICE


def test_autotune_rules():
    ba = K.Graph.generate(K.GEN_BA, seed=1, n=20000, attach=20).autotune(6)
    assert ba["orient"] == "degree" and ba["gran"] == "node" and ba["degree_max_out"] <= 64
    clg = K.Graph.generate(K.GEN_CHUNG_LU, seed=1, n=100000, m=600000, gamma=2.1, max_degree=5000.0)
    cl = clg.autotune(3)
    assert cl["orient"] == "degen" and cl["gran"] == "edge" and cl["degree_max_out"] > 64
    # the degree-order out-degree the plan reports is the oracle's
    G = K.Graph.generate(K.GEN_CHUNG_LU, seed=3, n=20000, m=80000, gamma=2.1, max_degree=900.0)
    rp, col = G.csr()
    orp, _ = O.orient(rp.astype(np.uint64), col)
    assert G.autotune(3)["degree_max_out"] == int(np.diff(orp).max())
    # above 64 out-neighbours at k >= 4 the plan is measured on the device:
    # without one the call fails loudly
    with pytest.raises(K.KCError) as ei:
        clg.autotune(6)
    assert ei.value.code == K.KC_ERR_DEVICE


def test_cli_usage_and_argument_errors(tmp_path):
    cli = os.path.join(ROOT, "kcgpu_amd", "bin", "kcgpu")
    r = subprocess.run([cli, "-h", "x"], capture_output=True, text=True)
    assert r.returncode == 0 and "Usage" in r.stderr
    g = tmp_path / "g.tsv"
    g.write_text("0 1\n")
    for args in (["-q", "x8"], ["-o", "full"], ["-m", "lkc"], ["-p", "vertex"]):
        r = subprocess.run([cli, "-g", str(g)] + args, capture_output=True, text=True)
        assert r.returncode != 0, args


def test_workload_mix_is_fixed():
    """The batched and serving workloads are defined by seeds alone: the same
    specs on every host (bench.py and the GPU tests rely on it)."""
    from kcgpu_amd.workloads import batch_mix, serve_mix, splitmix64
    a, b = batch_mix(64), batch_mix(64)
    assert a == b and len(a) == 64
    assert all(s["kind"] == ("ba" if i % 2 == 0 else "er") for i, s in enumerate(a))
    assert all(200 <= s["n"] <= 20000 and 4 <= s["k"] <= 7 for s in a)
    assert a[0]["seed"] == splitmix64(4 ^ 0) and a[5]["seed"] == splitmix64(4 ^ 5)
    assert all(3 <= s["attach"] <= 20 for s in a[::2])
    assert all(s["n"] * 4 // 2 <= s["m"] <= s["n"] * 40 // 2 for s in a[1::2])
    s1, s2 = serve_mix(40, seed=5), serve_mix(40, seed=5)
    assert s1 == s2 and [x["kind"] for x in s1[:4]] == ["er", "cl", "rmat", "ba"]
    # splitmix64 reference values (Steele, Lea, Flood 2014; state 0 and 1)
    assert splitmix64(0) == 0xE220A8397B1DCDAF and splitmix64(1) == 0x910A2DEC89025CC1


def _bench_module():
    import importlib.util
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py")
    spec = importlib.util.spec_from_file_location("kc_bench", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_sum_residues_of_chunked_counts(tmp_path):
    """bench.py --sum-residues: residue lines over several files add up per
    (orientation, granularity, k, stride); a residue seen twice with the same
    count counts once, missing residues are named, a residue with two
    different counts is an error, and a total of 2^64 or more is an overflow."""
    import json
    B = _bench_module()

    def line(o, off, c, n=4, ovf=False):
        return json.dumps({"orient": o, "gran": "edge", "k": 8, "stride": n, "offset": off, "count": c,
                           "overflow": ovf, "device_ms": 10.0}) + "\n"
    a, b = tmp_path / "a.log", tmp_path / "b.log"
    a.write_text("noise\n" + line("degen", 0, 5) + line("degen", 1, 7) + line("degree", 0, 2 ** 63))
    b.write_text(line("degen", 1, 7) + line("degen", 2, 11) + line("degen", 3, 13) + line("degree", 1, 2 ** 63)
                 + '{"orient": "degen", "gran": "edge", "k": 8, "residues": "0:4 of 4", "count": 36}\n')
    rows = {r["orient"]: r for r in B.sum_residues([str(a), str(b)])}
    assert rows["degen"]["count"] == 36 and rows["degen"]["missing"] == [] and not rows["degen"]["overflow"]
    assert rows["degen"]["have"] == 4 and rows["degen"]["device_s"] == 0.04
    assert rows["degree"]["missing"] == [2, 3] and rows["degree"]["overflow"]
    c = tmp_path / "c.log"
    c.write_text(line("degen", 0, 6))
    with pytest.raises(ValueError):
        B.sum_residues([str(a), str(c)])
