"""Learned near-field surrogate: Python side of the native MLP in libebem_hip.so.

The reference evaluates every element pair with the analytic routine
(functions/BEMenerg_analiticIntg/* via SAc/BEMenerg_coreSA_calcMatrixSubBlock.m);
this module adds the chosen learned form beside it:

* training pairs: near-field (test, field, lag) items -- pairs i != j whose
  centroid distance is at most 3h, lags inside the light-cone window of
  SAc/BEMenerg_coreSA_calcMatrixBlock.m:57-61 -- from generated icospheres
  (levels 2-4) and uniform screens (8-32 subdivisions), seeded;
* labels: the library's own FP64 analytic routine (P-front and S-front parts
  of the block, computed on the device by ebem_sur_add_samples);
* network, training (Adam), gradients and inference: hand-written BF16 MFMA
  kernels in the native library;
* use: ebem_assemble with PREC_SURROGATE predicts the near-field pairs whose
  features lie inside the training box; every other pair, and every pair when
  no surrogate is attached, goes through the analytic routine.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .mesh import SpaceMesh, icosphere, screen_uniform

N_IN = 24
N_OUT = 18
PARAMS = 230464
NORM = 132
LAYER_K = (64, 256, 256, 256, 256)
LAYER_N = (256, 256, 256, 256, 64)


def layer_offsets():
    """[(w_offset, b_offset, K, N)] of the flat parameter vector."""
    out, o = [], 0
    for K, N in zip(LAYER_K, LAYER_N):
        out.append((o, o + K * N, K, N))
        o += K * N + N
    assert o == PARAMS
    return out


def unpack(params: np.ndarray):
    """[(W (K, N), b (N,))] views of a flat parameter vector."""
    return [(params[w:w + K * N].reshape(K, N), params[b:b + N]) for w, b, K, N in layer_offsets()]


class Surrogate:
    """One native surrogate object (samples, parameters, optimiser state)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        s = C.c_void_p()
        _lib.check(self.lib.ebem_sur_create(C.byref(s), device))
        self.s = s

    def close(self):
        if getattr(self, "s", None):
            _lib.check(self.lib.ebem_sur_destroy(self.s))
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- samples ---------------------------------------------------------
    def add_samples(self, solver, test, field, lag) -> None:
        test = np.ascontiguousarray(test, np.int32)
        field = np.ascontiguousarray(field, np.int32)
        lag = np.ascontiguousarray(lag, np.int32)
        if not (test.shape == field.shape == lag.shape) or test.ndim != 1:
            raise ValueError("test, field and lag must be equal-length 1-D arrays")
        _lib.check(self.lib.ebem_sur_add_samples(self.s, solver.h, test.size, test.ctypes.data_as(_lib._ip),
                                                 field.ctypes.data_as(_lib._ip), lag.ctypes.data_as(_lib._ip)))

    def num_samples(self) -> int:
        n = C.c_int64()
        _lib.check(self.lib.ebem_sur_num_samples(self.s, C.byref(n)))
        return n.value

    def get_samples(self, first: int = 0, n: int | None = None):
        if n is None:
            n = self.num_samples() - first
        x = np.zeros((n, N_IN), np.float32)
        y = np.zeros((n, N_OUT), np.float32)
        _lib.check(self.lib.ebem_sur_get_samples(self.s, first, n, _lib.fptr(x), _lib.fptr(y)))
        return x, y

    # -- training ----------------------------------------------------------
    def init(self, ntrain: int, seed: int = 0) -> None:
        _lib.check(self.lib.ebem_sur_init(self.s, int(ntrain), int(seed)))

    def train(self, ntrain: int, epochs: int, batch: int = 2048, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        out = np.zeros(epochs)
        _lib.check(self.lib.ebem_sur_train(self.s, int(ntrain), int(epochs), int(batch), float(lr), int(seed),
                                           _lib.dptr(out)))
        return out

    def loss(self, first: int, n: int) -> float:
        v = C.c_double()
        _lib.check(self.lib.ebem_sur_loss(self.s, int(first), int(n), C.byref(v)))
        return v.value

    def grad(self, x: np.ndarray, y: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(PARAMS, np.float32)
        v = C.c_double()
        _lib.check(self.lib.ebem_sur_grad(self.s, x.shape[0], _lib.fptr(x), _lib.fptr(y), _lib.fptr(g), C.byref(v)))
        return g, v.value

    def predict(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], N_OUT), np.float32)
        _lib.check(self.lib.ebem_sur_predict(self.s, x.shape[0], _lib.fptr(x), _lib.fptr(y)))
        return y

    def params(self):
        p = np.zeros(PARAMS, np.float32)
        nm = np.zeros(NORM, np.float32)
        _lib.check(self.lib.ebem_sur_params(self.s, 0, _lib.fptr(p), _lib.fptr(nm)))
        return p, nm

    def set_params(self, p: np.ndarray, norm: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        nm = np.ascontiguousarray(norm, np.float32)
        if p.size != PARAMS or nm.size != NORM:
            raise ValueError(f"expected {PARAMS} parameters and {NORM} normalisation values")
        _lib.check(self.lib.ebem_sur_params(self.s, 1, _lib.fptr(p), _lib.fptr(nm)))

    def save(self, path: str) -> None:
        p, nm = self.params()
        np.savez(path, params=p, norm=nm)

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"], z["norm"])

    def attach(self, solver) -> None:
        _lib.check(self.lib.ebem_attach_surrogate(solver.h, self.s))

    @staticmethod
    def detach(solver) -> None:
        _lib.check(_lib.load().ebem_attach_surrogate(solver.h, None))


# ------------------------------------------------------------------ sampling --

def near_field_items(mesh: SpaceMesh, dt: float, cP: float, cS: float, L: int | None = None):
    """All near-field (test, field, lag) items: i != j, centroid distance <= 3h
    (h = longest edge), lag < L inside the light-cone window of
    SAc/BEMenerg_coreSA_calcMatrixBlock.m:57-61 (same expressions as the library)."""
    h = mesh.lMax
    if L is None:
        L = int(math.ceil(5.0 * h / (cS * dt))) + 4
    N = mesh.numberTriangles
    c, mx = mesh.center, mesh.maxL
    I, J = [], []
    for r0 in range(0, N, 512):
        rows = np.arange(r0, min(N, r0 + 512))
        d = np.sqrt(((c[None, :, :] - c[rows, None, :]) ** 2).sum(-1))
        ii, jj = np.nonzero((d <= 3.0 * h) & (rows[:, None] != np.arange(N)[None, :]))
        I.append(rows[ii])
        J.append(jj)
    I = np.concatenate(I).astype(np.int32)
    J = np.concatenate(J).astype(np.int32)
    d = np.sqrt(((c[J] - c[I]) ** 2).sum(-1))
    dmin = d - mx[J] - mx[I]
    dmax = d + mx[J] + mx[I]
    lags = np.arange(L)
    ok = (((lags - 1) * cS * dt)[None, :] < dmax[:, None]) & (((lags + 1) * cP * dt)[None, :] >= dmin[:, None])
    pi, lag = np.nonzero(ok)
    return I[pi], J[pi], lag.astype(np.int32)


@dataclass
class TrainConfig:
    name: str
    mesh: SpaceMesh
    dt: float


def training_configs():
    """Generated meshes and time steps of the training set (lambda = mu = rho = 1).

    Icospheres of levels 2-4 and screens of 8, 16, 24, 32 subdivisions; time
    steps chosen so that dt c_S / h spans 0.1-1.4 and includes the steps of the
    benchmark problems (sphere level 3 at dt = 1/32, screens at dt = 1/16 and
    1/32; a 64 x 64 screen at dt = 1/32 has the features of the 32 x 32 screen
    at dt = 1/16)."""
    cfgs = []
    for lev in (2, 3, 4):
        m = icosphere(lev)
        h = m.lMax
        for r in (0.1, 0.19, 0.4, 0.8):
            cfgs.append(TrainConfig(f"ico{lev}_r{r}", m, r * h))
    m3 = icosphere(3)
    cfgs.append(TrainConfig("ico3_dt1/32", m3, 1.0 / 32))
    for n in (8, 16, 24, 32):
        m = screen_uniform(n)
        for dt in (1.0 / 16, 1.0 / 32):
            cfgs.append(TrainConfig(f"screen{n}_dt1/{round(1 / dt)}", m, dt))
    return cfgs


def build_training_set(sur: Surrogate, n_total: int, seed: int = 0, configs=None, device: int = 0,
                       held_frac: float = 0.0, log=None):
    """Append about n_total seeded near-field samples (equal quota per config),
    then, disjoint from them, about held_frac * n_total held-out samples.
    Returns (training samples, held-out samples): the training set is the
    prefix [0, n_train) of the surrogate's samples."""
    from .api import PbParam, Solver
    configs = training_configs() if configs is None else configs
    rng = np.random.default_rng(seed)
    quota = int(math.ceil(n_total / len(configs)))
    hquota = int(math.ceil(held_frac * n_total / len(configs)))
    picks = []
    for cfg in configs:
        pb = PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=cfg.dt * 64, nT=64)
        i, j, g = near_field_items(cfg.mesh, cfg.dt, pb.velP, pb.velS)
        sel = rng.permutation(i.size)[:quota + hquota]
        tr, he = np.sort(sel[:quota]), np.sort(sel[quota:])
        picks.append((cfg, pb, (i[tr], j[tr], g[tr]), (i[he], j[he], g[he])))
        if log:
            log(f"surrogate samples: {cfg.name}: {tr.size} + {he.size} held out of {i.size}")
    counts = [0, 0]
    for part in (0, 1):
        for cfg, pb, trs, hes in picks:
            ii, jj, gg = (trs, hes)[part]
            if ii.size == 0:
                continue
            with Solver(cfg.mesh, pb, 7, device=device) as s:
                sur.add_samples(s, ii, jj, gg)
            counts[part] += int(ii.size)
    return counts[0], counts[1]
