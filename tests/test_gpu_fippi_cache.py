"""StoreCache / PreloadCache round trip on the GPU path: the records a
geometry computes are written in the reference's file format, and a
fresh process that preloads the file builds the same geometry's table
from it (bit-identical records and matrix; the cache does not grow, so
no record is computed).  In one process the cache is shared by every
geometry, as the reference's GlobalFIPPICache is."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG, ROOT

pytestmark = pytest.mark.gpu

CHILD = r"""
import sys, numpy as np
sys.path.insert(0, {pkg!r}); sys.path.insert(0, {root!r})
import scuffx
from oracle import oracle as O
n = scuffx.PreloadCache({cache!r})
print("preloaded", n, flush=True)
V, T = O.icosphere(2)
G = scuffx.RWGGeometry()
G.AddMesh(V, T, "CONST_EPS_4+0.1i")
M = G.AssembleBEMMatrix(0.9).numpy()
R = G.FIPPIRecords()
np.save({out_m!r}, M)
np.save({out_r!r}, R)
print("cache", n, scuffx.CacheSize(), "fippi_ms", G.AssemblyStats()["fippi_ms"], flush=True)
"""


def test_store_then_preload_in_fresh_process(sx, O, tmp_path):
    print("fippi cache round trip", flush=True)
    V, T = O.icosphere(2)
    G = sx.RWGGeometry()
    G.AddMesh(V, T, "CONST_EPS_4+0.1i")
    M = G.AssembleBEMMatrix(0.9).numpy()
    R = G.FIPPIRecords()
    t_build = G.AssemblyStats()["fippi_ms"]
    cache = str(tmp_path / "sphere.cache")
    nst = sx.StoreCache(cache)
    assert nst >= len(R)
    assert (os.path.getsize(cache) - 11) % 352 == 0
    out_m, out_r = str(tmp_path / "m.npy"), str(tmp_path / "r.npy")
    code = CHILD.format(pkg=PKG, root=ROOT, cache=cache, out_m=out_m, out_r=out_r)
    p = subprocess.run([sys.executable, "-u", "-c", code], capture_output=True, text=True, timeout=200)
    print(p.stdout, p.stderr[-2000:], flush=True)
    assert p.returncode == 0
    assert f"preloaded {nst}" in p.stdout
    # every record of the new table came from the file: the cache did not grow
    assert f"cache {nst} {nst} " in p.stdout
    assert np.array_equal(np.load(out_r), R)
    assert np.array_equal(np.load(out_m), M)
    t_pre = float(p.stdout.split("fippi_ms")[1].split()[0])
    print(f"  FIPPI table: {t_build:.1f} ms in this process, {t_pre:.1f} ms from the file", flush=True)
