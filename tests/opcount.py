"""Algorithmic FP64 operation count per (pair, level) item of the assembly
kernel, from the op-counting host build of the device math
(tests/native/opcount.cpp), on a uniform random sample of the items of a
workload.  Run:  python tests/opcount.py [nsample]   (sphereUniform_3, T=4, nT=128).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
_dp = C.POINTER(C.c_double)


def items_of(o, L):
    """First level and item count of every pair under the library's pruned
    lag windows (the items the assembly kernel evaluates)."""
    from hostasm import exact_windows
    lo, hi = exact_windows(o, L)
    mlo = np.maximum(1, lo - 1)
    n = np.where(hi >= lo, hi + 1 - mlo + 1, 0)
    return mlo, n


def count(mesh, Tfin, nT, L, nsample, seed=0, nGH=7):
    from oracle import Oracle
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native")], check=True)
    lib = C.CDLL(os.path.join(HERE, "native", "libopcount.so"))
    lib.opcount_item.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double]
    lib.opcount_item.restype = C.c_longlong
    o = Oracle(mesh.coordinates, mesh.triangles, Tfin=Tfin, nT=nT, nGH=nGH)
    mlo, n = items_of(o, L)
    flat = n.ravel()
    off = np.concatenate([[0], np.cumsum(flat)])
    total = int(off[-1])
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, total, size=nsample)
    pair = np.searchsorted(off, picks, side="right") - 1
    N = mesh.numberTriangles
    recs = {}
    ops = np.zeros(nsample)
    for k, (it, p) in enumerate(zip(picks, pair)):
        i, j = divmod(int(p), N)
        m = int(mlo.ravel()[p] + (it - off[p]))
        if j not in recs:
            T = o.coords[o.tri[j]]
            recs[j] = o.field_record(T, o.normal[j])
        V = o.coords[o.tri[i]]
        gh = np.zeros(128)
        nodes = o.ghn @ V
        gh[0:nGH], gh[32:32 + nGH], gh[64:64 + nGH] = nodes[:, 0], nodes[:, 1], nodes[:, 2]
        gh[96:96 + nGH] = o.area[i] * o.ghw
        ops[k] = lib.opcount_item(recs[j].ctypes.data_as(_dp), gh.ctypes.data_as(_dp), nGH, m * o.dt, o.velP,
                                  o.velS)
    return total, ops.mean(), ops.std() / np.sqrt(nsample)


if __name__ == "__main__":
    from elastodynamicebem_amd.mesh import read_mesh
    ns = int(sys.argv[1]) if len(sys.argv) > 1 else 20000
    m = read_mesh(os.path.join(HERE, "golden", "meshes", "sphereUniform_3.mesh"))
    total, mean, se = count(m, 4.0, 128, 66, ns)
    print(f"items {total}  flops/item {mean:.1f} +- {se:.1f} (n={ns})")
