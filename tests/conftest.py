"""Shared pytest set-up: the `gpu` marker, import paths and fixtures.

`-m "not gpu"` runs the oracle pins and host-side checks on any machine;
`-m gpu` runs the parity tests that call libscuffx.so on a GPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "scuff-em_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (calls libscuffx.so kernels)")


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def sx():
    import scuffx
    return scuffx


@pytest.fixture(autouse=True)
def _fresh_fippi_cache(request):
    """The product's FIPPI cache is process-wide, as the reference's is;
    the oracle's lives with one geometry.  Parity tests start every test
    from an empty cache so both sides reuse records only within the test's
    own geometry."""
    if request.node.get_closest_marker("gpu") is not None:
        import scuffx
        scuffx.ClearCache()
    yield


def rel_fro(A, B):
    """Relative Frobenius distance ||A-B|| / ||B||."""
    return float(np.linalg.norm(np.asarray(A) - np.asarray(B)) / np.linalg.norm(np.asarray(B)))


def read_msh(path):
    """Vertices (file order) and triangles (0-based) of a GMSH 2.x ASCII
    mesh: every node is kept, every type-2 element is a panel."""
    lines = open(path).read().split("\n")
    nodes, tris, i = {}, [], 0
    while i < len(lines):
        tag = lines[i].strip()
        if tag == "$Nodes":
            n = int(lines[i + 1])
            for ln in lines[i + 2:i + 2 + n]:
                a = ln.split()
                nodes[int(a[0])] = [float(x) for x in a[1:4]]
            i += n + 2
        elif tag == "$Elements":
            n = int(lines[i + 1])
            for ln in lines[i + 2:i + 2 + n]:
                a = [int(x) for x in ln.split()]
                if a[1] == 2:
                    tris.append(a[3 + a[2]:6 + a[2]])
            i += n + 2
        else:
            i += 1
    keys = list(nodes)
    pos = {k: j for j, k in enumerate(keys)}
    V = np.array([nodes[k] for k in keys])
    T = np.array([[pos[t] for t in tri] for tri in tris], dtype=np.int32)
    return V, T
