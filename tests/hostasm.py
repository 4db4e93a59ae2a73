"""Host replay of the GPU assembly (items W_m -> second differences -> blocks)
using a host build of the device math; used to study numerics on the CPU."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)


def load(fma=False):
    lib = C.CDLL(os.path.join(HERE, "native", "libsa_host_fma.so" if fma else "libsa_host.so"))
    lib.sa_item.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double, _dp, C.c_int]
    return lib


def windows(o, L):
    c, mx, dt = o.center, o.maxL, o.dt
    d = np.sqrt(((c[:, None, :] - c[None, :, :]) ** 2).sum(-1))
    dmin = d - mx[None, :] - mx[:, None]
    dmax = d + mx[None, :] + mx[:, None]
    lo = np.full(d.shape, L, dtype=np.int64)
    hi = np.full(d.shape, -1, dtype=np.int64)
    for l in range(L):
        ok = ((l - 1) * o.velS * dt < dmax) & ((l + 1) * o.velP * dt >= dmin)
        lo = np.where(ok & (lo == L), l, lo)
        hi = np.where(ok, l, hi)
    return lo, hi


def assemble(o, lib, L, rows=None, lags=None, fp32_far=None):
    """Dict lag -> dense E (3N x 3N) restricted to `rows`."""
    N = o.N
    lo, hi = windows(o, L)
    rows = range(N) if rows is None else rows
    lags = range(L) if lags is None else lags
    E = {l: np.zeros((3 * N, 3 * N)) for l in lags}
    recs = [o.field_record(o.coords[o.tri[j]], o.normal[j]) for j in range(N)]
    qq = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    inv = 1.0 / (4 * np.pi * o.rho)
    for i in rows:
        V = o.coords[o.tri[i]]
        gh = np.zeros(128)
        # same operation order as the library's host set-up (and the oracle)
        nodes = (o.ghn[:, 0:1] * V[0][None, :] + o.ghn[:, 1:2] * V[1][None, :]) + o.ghn[:, 2:3] * V[2][None, :]
        n = o.nGH
        gh[0:n], gh[32:32 + n], gh[64:64 + n] = nodes[:, 0], nodes[:, 1], nodes[:, 2]
        gh[96:96 + n] = o.area[i] * o.ghw
        for j in range(N):
            a, b = lo[i, j], hi[i, j]
            if a > b:
                continue
            W = {}
            f32 = int(fp32_far is not None and np.linalg.norm(o.center[i] - o.center[j]) > fp32_far)
            for m in range(max(1, a - 1), b + 2):
                w = np.zeros(6)
                lib.sa_item(recs[j].ctypes.data_as(_dp), gh.ctypes.data_as(_dp), n, m * o.dt, o.velP, o.velS,
                            w.ctypes.data_as(_dp), f32)
                W[m] = w
            for l in lags:
                if not (a <= l <= b):
                    continue
                e = 1.0 * W[l + 1]
                if l >= 1:
                    e = e + (-2.0) * W[l]
                if l >= 2:
                    e = e + W[l - 1]
                e = e * inv
                e[np.abs(e) < 1e-14] = 0
                blk = np.zeros((3, 3))
                for q, (r, c) in enumerate(qq):
                    blk[r, c] = blk[c, r] = e[q]
                E[l][3 * i:3 * i + 3, 3 * j:3 * j + 3] = blk
    return E


def exact_windows(o, L):
    """The library's pruned lag windows: the reference's cull window of
    calcMatrixBlock.m:57-61 cut to the lags whose block is not zero in exact
    arithmetic (after the S front: every outer node of the test triangle sees
    the field triangle inside its S disk at level l-1; before the P front:
    level l+1 cannot reach the field triangle from any node)."""
    lo, hi = windows(o, L)
    N, dt = o.N, o.dt
    V = o.coords[o.tri]
    X = (o.ghn[None, :, 0:1] * V[:, None, 0, :] + o.ghn[None, :, 1:2] * V[:, None, 1, :]) + \
        o.ghn[None, :, 2:3] * V[:, None, 2, :]                      # N, nGH, 3
    RF = np.sqrt(((V - o.center[:, None, :]) ** 2).sum(-1)).max(1)
    lo2, hi2 = lo.copy(), hi.copy()
    for i in range(N):
        dmax = np.sqrt(((X[i][None, :, None, :] - V[:, None, :, :]) ** 2).sum(-1)).max(axis=(1, 2))
        dlo = np.maximum(0.0, np.sqrt(((X[i][None, :, :] - o.center[:, None, :]) ** 2).sum(-1)).min(1) - RF)
        ms = np.floor(dmax * (1 + 1e-12) / (o.velS * dt)) + 1
        ml = np.ceil(dlo * (1 - 1e-12) / (o.velP * dt) - 1)
        ok = lo[i] <= hi[i]
        hi2[i] = np.where(ok, np.minimum(hi[i], ms), hi[i]).astype(np.int64)
        lo2[i] = np.where(ok, np.maximum(lo[i], ml), lo[i]).astype(np.int64)
    return lo2, hi2
