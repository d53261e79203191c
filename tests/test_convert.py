"""CPU tests of the format converters (-m mtx-bel, tsv-bel, tsv-mtx, bel-mtx,
txt-bel; include/FIleReader.cuh:480-933) and of the oriented-CSR cache file's
validation.  Expected outputs are written out by hand from the reference's
converter semantics for each tiny input.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import kcgpu_amd as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "kcgpu_amd", "bin", "kcgpu")


def _bel(path):
    b = open(path, "rb").read()
    assert len(b) % 24 == 0
    return [struct.unpack_from("<QQQ", b, i) for i in range(0, len(b), 24)]


def _edges(G):
    rp, col = G.csr()
    return sorted((u, int(v)) for u in range(len(rp) - 1) for v in col[rp[u]:rp[u + 1]])


def test_mtx_to_bel_symmetric(tmp_path):
    src, dst = tmp_path / "a.mtx", tmp_path / "a.bel"
    src.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% a comment\n4 4 4\n2 1\n3 1\n3 3\n4 2\n")
    K.convert("mtx-bel", src, dst)
    # pairs (row, col) plus (col, row) for row != col, sorted; record = (dst=col, src=row, 0)
    pairs = sorted([(2, 1), (1, 2), (3, 1), (1, 3), (3, 3), (4, 2), (2, 4)])
    assert _bel(dst) == [(c, r, 0) for r, c in pairs]
    # the BEL file loads into the same simple graph as the MTX file
    assert _edges(K.Graph.load(dst)) == _edges(K.Graph.load(src))


def test_mtx_to_bel_reference_flags_and_values(tmp_path):
    src, dst = tmp_path / "a.mtx", tmp_path / "a.bel"
    src.write_text("%%MatrixMarket matrix coordinate real symmetric\n% made by hand\n3 3 2\n2 1 0.5\n3 2 7\n")
    K.convert("mtx-bel", src, dst)
    assert _bel(dst) == [(2, 1, 0), (1, 2, 0), (3, 2, 0), (2, 3, 0)]
    # the reference re-reads the flags on every '%' line: the comment clears "symmetric"
    K.convert("mtx-bel", src, dst, reference_flags=True)
    assert _bel(dst) == [(1, 2, 0), (2, 3, 0)]


def test_tsv_to_bel_and_mtx(tmp_path):
    src = tmp_path / "a.tsv"
    # "dst src weight" triples; reading stops at the first incomplete triple
    src.write_text("2 1 5\n3 1 1\n1 2 9\n4 3 0\nx 9 9\n5 5 5\n")
    K.convert("tsv-bel", src, tmp_path / "a.bel")
    assert _bel(tmp_path / "a.bel") == [(2, 1, 0), (3, 1, 0), (1, 2, 0), (4, 3, 0)]
    K.convert("tsv-mtx", src, tmp_path / "a.mtx")
    # n = src of the last edge; lines "src dst weight"
    assert (tmp_path / "a.mtx").read_text() == ("%%MatrixMarket matrix coordinate general\n3 3 4\n"
                                                 "1 2 5\n1 3 1\n2 1 9\n3 4 0\n")


def test_bel_to_mtx(tmp_path):
    src = tmp_path / "a.bel"
    recs = [(2, 1, 7), (5, 1, (1 << 32) + 3), (1, 4, 0xFFFFFFFF)]
    src.write_bytes(b"".join(struct.pack("<QQQ", *r) for r in recs) + b"\x01\x02")  # partial record dropped
    K.convert("bel-mtx", src, tmp_path / "a.mtx")
    # weight = low 32 bits of the u64 read as int
    assert (tmp_path / "a.mtx").read_text() == ("%%MatrixMarket matrix coordinate general\n4 4 3\n"
                                                 "1 2 7\n1 5 3\n4 1 -1\n")


def test_txt_to_bel(tmp_path):
    src = tmp_path / "a.txt"
    src.write_text("% header\n# snap comment\n1 2\n2 1\n\n3 3\n0 4 17\n2 1\n")
    K.convert("txt-bel", src, tmp_path / "a.bel")
    pairs = sorted({(1, 2), (2, 1), (0, 4), (4, 0)})
    assert _bel(tmp_path / "a.bel") == [(b, a, 0) for a, b in pairs]


def test_round_trip_txt_bel_mtx(tmp_path):
    rng = np.random.default_rng(3)
    e = rng.integers(0, 200, size=(1500, 2))
    (tmp_path / "g.txt").write_text("".join(f"{a} {b}\n" for a, b in e))
    K.convert("txt-bel", tmp_path / "g.txt", tmp_path / "g.bel")
    K.convert("bel-mtx", tmp_path / "g.bel", tmp_path / "g.mtx")
    ref = K.Graph.from_edges(e[:, 0].astype(np.uint32), e[:, 1].astype(np.uint32))
    assert _edges(K.Graph.load(tmp_path / "g.bel")) == _edges(ref)
    assert _edges(K.Graph.load(tmp_path / "g.mtx")) == _edges(ref)


def test_converter_errors(tmp_path):
    with pytest.raises(K.KCError) as ei:
        K.convert("mtx-bel", tmp_path / "missing.mtx", tmp_path / "o.bel")
    assert ei.value.code == K.KC_ERR_IO
    (tmp_path / "e.tsv").write_text("% only a comment\n")
    with pytest.raises(K.KCError) as ei:  # no edges: the reference reads fileEdges[-1]
        K.convert("tsv-mtx", tmp_path / "e.tsv", tmp_path / "o.mtx")
    assert ei.value.code == K.KC_ERR_FORMAT
    (tmp_path / "r.mtx").write_text("%%MatrixMarket matrix coordinate pattern general\n3 4 1\n1 2\n")
    with pytest.raises(K.KCError) as ei:
        K.convert("mtx-bel", tmp_path / "r.mtx", tmp_path / "o.bel")
    assert ei.value.code == K.KC_ERR_FORMAT
    (tmp_path / "s.mtx").write_text("%%MatrixMarket matrix coordinate pattern general\n3 3 3\n1 2\n")
    with pytest.raises(K.KCError) as ei:  # fewer entries than announced
        K.convert("mtx-bel", tmp_path / "s.mtx", tmp_path / "o.bel")
    assert ei.value.code == K.KC_ERR_FORMAT
    with pytest.raises(K.KCError) as ei:
        K.convert(9, tmp_path / "s.mtx", tmp_path / "o.bel")
    assert ei.value.code == K.KC_ERR_ARG


def test_cli_conversion(tmp_path):
    (tmp_path / "a.txt").write_text("1 2\n2 3\n")
    r = subprocess.run([CLI, "-m", "txt-bel", "-g", str(tmp_path / "a.txt"), "-r", str(tmp_path / "a.bel")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert _bel(tmp_path / "a.bel") == [(2, 1, 0), (1, 2, 0), (3, 2, 0), (2, 3, 0)]
    r = subprocess.run([CLI, "-m", "tsv-mtx", "-g", str(tmp_path / "a.txt")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "-r" in r.stderr


# ---- oriented-CSR cache file ----------------------------------------------------
def _fnv(words, h=14695981039346656037):
    for w in np.asarray(words, dtype="<u4").tobytes():
        h = ((h ^ w) * 1099511628211) & (2**64 - 1)
    return h


def _cache_bytes(rp, col, key, mode=1, max_out=None, max_core=2, checksum=None):
    rp, col, key = (np.asarray(a, dtype="<u4") for a in (rp, col, key))
    n, m = len(rp) - 1, len(col)
    if max_out is None:
        max_out = int(np.diff(rp).max()) if n else 0
    if checksum is None:
        checksum = _fnv(key, _fnv(col, _fnv(rp)))
    hdr = b"KCOCSR1\0" + struct.pack("<IIQIIQ", n, mode, m, max_out, max_core, checksum)
    return hdr + rp.tobytes() + col.tobytes() + key.tobytes()


def test_cache_file_validation(tmp_path):
    p = tmp_path / "g.kco"
    # triangle 0->1, 0->2, 1->2 by (key, id)
    good = dict(rp=[0, 2, 3, 3], col=[1, 2, 2], key=[0, 0, 0])
    bad = {
        "magic": b"XXOCSR1\0" + _cache_bytes(**good)[8:],
        "checksum": _cache_bytes(**good, checksum=1),
        "cycle": _cache_bytes(rp=[0, 1, 2, 3], col=[1, 2, 0], key=[0, 0, 0]),
        "unsorted": _cache_bytes(rp=[0, 2, 3, 3], col=[2, 1, 2], key=[0, 0, 0]),
        "max_out": _cache_bytes(**good, max_out=5),
        "truncated": _cache_bytes(**good)[:-4],
        "trailing": _cache_bytes(**good) + b"\0",
        "mode": _cache_bytes(**good, mode=7),
    }
    for name, b in bad.items():
        p.write_bytes(b)
        with pytest.raises(K.KCError) as ei:
            K.Oriented.load(p)
        assert ei.value.code == K.KC_ERR_FORMAT, name
    # a well-formed file passes every check and only then needs the device
    p.write_bytes(_cache_bytes(**good))
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        with pytest.raises(K.KCError) as ei:
            K.Oriented.load(p)
        assert ei.value.code == K.KC_ERR_DEVICE
