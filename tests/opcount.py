"""Algorithmic FP64 operation count per (pair, level) item of the assembly
(items outside the kP and between-front ranges; also with the static outer
nodes whose value an earlier level holds left out, as the node-task path
evaluates them), per kP range
of k_levels_kp and per between-front range of k_levels_bf, from the op-counting host build of the device math
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
    """Uniform sample of the items: mean flops of the per-level items (k_levels)
    and of the kP ranges (k_levels_kp, per range, writes included), with the
    share of items inside kP ranges."""
    from hostasm import load
    from oracle import Oracle
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native")], check=True)
    lib = C.CDLL(os.path.join(HERE, "native", "libopcount.so"))
    lib.opcount_item.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double]
    lib.opcount_item.restype = C.c_longlong
    lib.opcount_item_reuse.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    lib.opcount_item_reuse.restype = C.c_longlong
    lib.opcount_kp.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    lib.opcount_kp.restype = C.c_longlong
    lib.opcount_bf.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
    lib.opcount_bf.restype = C.c_longlong
    host = load()
    host.sa_kp_range.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    o = Oracle(mesh.coordinates, mesh.triangles, Tfin=Tfin, nT=nT, nGH=nGH)
    from hostasm import exact_windows
    lo, hi = exact_windows(o, L)
    mlo, n = items_of(o, L)
    flat = n.ravel()
    off = np.concatenate([[0], np.cumsum(flat)])
    total = int(off[-1])
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, total, size=nsample)
    pair = np.searchsorted(off, picks, side="right") - 1
    N = mesh.numberTriangles
    recs = {}
    lev_ops, lev_reuse, kp_ops, kp_w, bf_ops, bf_w = [], [], [], [], [], []
    for it, p in zip(picks, pair):
        i, j = divmod(int(p), N)
        m = int(mlo.ravel()[p] + (it - off[p]))
        if j not in recs:
            T = o.coords[o.tri[j]]
            recs[j] = o.field_record(T, o.normal[j])
        V = o.coords[o.tri[i]]
        gh = np.zeros(128)
        nodes = (o.ghn[:, 0:1] * V[0][None, :] + o.ghn[:, 1:2] * V[1][None, :]) + o.ghn[:, 2:3] * V[2][None, :]
        gh[0:nGH], gh[32:32 + nGH], gh[64:64 + nGH] = nodes[:, 0], nodes[:, 1], nodes[:, 2]
        gh[96:96 + nGH] = o.area[i] * o.ghw
        vF = np.ascontiguousarray(o.coords[o.tri[j]].ravel())
        ka, kb, sa, sb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        host.sa_kp_range(recs[j].ctypes.data_as(_dp), gh.ctypes.data_as(_dp), vF.ctypes.data_as(_dp), nGH, o.velP,
                         o.velS, o.dt, int(lo[i, j]), int(hi[i, j]), C.byref(ka), C.byref(kb), C.byref(sa),
                         C.byref(sb))
        if ka.value <= m <= kb.value:
            nl = kb.value - ka.value + 1
            c = lib.opcount_kp(recs[j].ctypes.data_as(_dp), gh.ctypes.data_as(_dp), nGH, ka.value * o.dt,
                               kb.value * o.dt, o.velP, o.velS)
            kp_ops.append(c + 13 * nl)
            kp_w.append(1.0 / nl)
        elif sa.value <= m <= sb.value:
            nl = sb.value - sa.value + 1
            c = lib.opcount_bf(recs[j].ctypes.data_as(_dp), gh.ctypes.data_as(_dp), nGH, sa.value, nl, o.dt,
                               o.velP, o.velS)
            bf_ops.append(c)
            bf_w.append(1.0 / nl)
        else:
            lev_ops.append(lib.opcount_item(recs[j].ctypes.data_as(_dp), gh.ctypes.data_as(_dp), nGH, m * o.dt,
                                            o.velP, o.velS))
            tprev = (m - 1) * o.dt if m > mlo.ravel()[p] else -1.0
            lev_reuse.append(lib.opcount_item_reuse(recs[j].ctypes.data_as(_dp), gh.ctypes.data_as(_dp), nGH,
                                                    m * o.dt, tprev, o.velP, o.velS))
    lev = np.array(lev_ops, float)
    levr = np.array(lev_reuse, float)
    kpo, kpw = np.array(kp_ops, float), np.array(kp_w)
    bfo, bfw = np.array(bf_ops, float), np.array(bf_w)
    return dict(items=total, kp_share=len(kpo) / nsample, bf_share=len(bfo) / nsample, level_flops=lev.mean(),
                level_flops_se=lev.std() / np.sqrt(len(lev)), level_flops_reuse=levr.mean(),
                level_flops_reuse_se=levr.std() / np.sqrt(len(levr)),
                kp_range_flops=float((kpo * kpw).sum() / kpw.sum()) if len(kpo) else 0.0,
                bf_level_flops=float(bfo.sum() / (bfw > 0).sum() / (1.0 / bfw).mean()) if len(bfo) else 0.0,
                bf_range_flops=float((bfo * bfw).sum() / bfw.sum()) if len(bfo) else 0.0)


if __name__ == "__main__":
    from elastodynamicebem_amd.mesh import read_mesh
    ns = int(sys.argv[1]) if len(sys.argv) > 1 else 20000
    m = read_mesh(os.path.join(HERE, "golden", "meshes", "sphereUniform_3.mesh"))
    r = count(m, 4.0, 128, 66, ns)
    print(f"items {r['items']}  kP share {r['kp_share']:.4f}  between-front share {r['bf_share']:.4f}  "
          f"per-level flops/item {r['level_flops']:.1f} +- {r['level_flops_se']:.1f} (static nodes reused: "
          f"{r['level_flops_reuse']:.1f} +- {r['level_flops_reuse_se']:.1f})  flops per kP range "
          f"{r['kp_range_flops']:.0f}, per between-front range {r['bf_range_flops']:.0f} (n={ns})")
