import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP kernels through the C ABI")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "cliques.json")) as f:
        d = json.load(f)
    for g in d["graphs"]:
        g["counts"] = {int(k): v for k, v in g["counts"].items()}
    return d["graphs"]


@pytest.fixture(scope="session")
def near_cliques():
    """complements of sparse graphs with exact counts (tests/golden/make_golden.py)"""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "cliques.json")) as f:
        d = json.load(f)
    for g in d["near_cliques"]:
        g["counts"] = {int(k): v for k, v in g["counts"].items()}
    return d["near_cliques"]


def complement_edges(n, h_edges):
    """edge list (src, dst) of the complement of the graph on n vertices with edges h_edges"""
    import numpy as np
    a, b = np.triu_indices(n, 1)
    keep = np.ones(a.size, dtype=bool)
    if len(h_edges):
        h = np.array(h_edges, dtype=np.int64).reshape(-1, 2)
        lo, hi = h.min(axis=1), h.max(axis=1)
        # index of pair (i, j), i < j, in triu_indices order
        idx = lo * n - lo * (lo + 1) // 2 + (hi - lo - 1)
        keep[idx] = False
    return a[keep].astype(np.uint32), b[keep].astype(np.uint32)
