"""CPU oracle of the near-field surrogate MLP (test infrastructure only).

NumPy FP64 restatement of the network the native library trains and runs
(24 -> 256 -> 256 -> 256 -> 256 -> 18, exact GELU, padded parameter layout of
include/ebem.h), with its forward pass, MSE loss and backward pass.  With
``emulate=True`` it rounds to BF16 at the same points as the device kernels
(standardised inputs, weights, hidden activations, output and hidden
gradients before each product), so that what remains between the two is
FP32-vs-FP64 accumulation; with ``emulate=False`` it is plain FP64 and is
pinned by central finite differences in tests/test_surrogate.py.

Nothing in the shipped library imports this file.  The reference has no
learned component, so there is no reference file to cite for the network; the
pair features and labels it is trained on are pinned against the SA oracle.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

N_IN, N_OUT = 24, 18
LAYER_K = (64, 256, 256, 256, 256)
LAYER_N = (256, 256, 256, 256, 64)
PARAMS = sum(k * n + n for k, n in zip(LAYER_K, LAYER_N))
NORM = 2 * N_IN + 2 * N_OUT + 2 * N_IN


def bf16(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def layers(params: np.ndarray):
    out, o = [], 0
    for K, N in zip(LAYER_K, LAYER_N):
        out.append((params[o:o + K * N].reshape(K, N), params[o + K * N:o + K * N + N], o, o + K * N))
        o += K * N + N
    return out


def standardise(x: np.ndarray, y: np.ndarray, norm: np.ndarray):
    """Float32 standardisation exactly as the gather kernel: (x - mean) * (1/std), (y - mean) / std."""
    norm = np.asarray(norm, np.float32)
    xs = (np.asarray(x, np.float32) - norm[:N_IN]) * norm[N_IN:2 * N_IN]
    ym, ysd = norm[2 * N_IN:2 * N_IN + N_OUT], norm[2 * N_IN + N_OUT:2 * N_IN + 2 * N_OUT]
    ys = None if y is None else (np.asarray(y, np.float32) - ym) / ysd
    return xs.astype(np.float32), None if ys is None else ys.astype(np.float32)


def forward(params, norm, x, emulate=True):
    """Standardised outputs (n x 18, float64) and the cache of the backward pass."""
    xs, _ = standardise(x, None, norm)
    n = xs.shape[0]
    a = np.zeros((n, LAYER_K[0]), np.float64)
    a[:, :N_IN] = bf16(xs) if emulate else xs
    cache = []
    L = layers(np.asarray(params, np.float32))
    for l, (W, b, _, _) in enumerate(L):
        Wd = (bf16(W) if emulate else W).astype(np.float64)
        z = a @ Wd + b.astype(np.float64)
        cache.append((a, z))
        if l < len(L) - 1:
            g = gelu(z)
            a = bf16(g.astype(np.float32)).astype(np.float64) if emulate else g
        else:
            a = z
    return a[:, :N_OUT], cache


def loss_grad(params, norm, x, y, emulate=True):
    """MSE over standardised labels (mean over n x 18) and its gradient (PARAMS,)."""
    params = np.asarray(params, np.float32)
    _, ys = standardise(x, y, norm)
    out, cache = forward(params, norm, x, emulate)
    n = out.shape[0]
    d = out - ys.astype(np.float64)
    loss = float((d * d).sum() / (n * N_OUT))
    g = np.zeros(PARAMS, np.float64)
    dz = np.zeros((n, LAYER_N[-1]))
    dz[:, :N_OUT] = 2.0 * d / (n * N_OUT)
    if emulate:
        dz = dz.astype(np.float32).astype(np.float64)
    L = layers(params)
    for l in range(len(L) - 1, -1, -1):
        W, b, wo, bo = L[l]
        a, _ = cache[l]
        dzb = bf16(dz.astype(np.float32)).astype(np.float64) if emulate else dz
        g[wo:wo + W.size] = (a.T @ dzb).ravel()
        g[bo:bo + b.size] = dz.sum(0)
        if l > 0:
            Wd = (bf16(W) if emulate else W).astype(np.float64)
            da = dzb @ Wd.T
            zprev = cache[l - 1][1]
            dz = da * gelu_d(zprev)
            if emulate:
                dz = dz.astype(np.float32).astype(np.float64)
    return loss, g


def predict(params, norm, x, emulate=True):
    """Unstandardised outputs (n x 18)."""
    out, _ = forward(params, norm, x, emulate)
    norm = np.asarray(norm, np.float64)
    return out * norm[2 * N_IN + N_OUT:2 * N_IN + 2 * N_OUT] + norm[2 * N_IN:2 * N_IN + N_OUT]


def random_params(seed=0, scale=1.0):
    """He-normal weights on the real (unpadded) rows/columns, small random biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    for l, (W, b, wo, bo) in enumerate(layers(p)):
        K, N = W.shape
        kin = N_IN if l == 0 else K
        nout = N_OUT if l == len(LAYER_K) - 1 else N
        Wr = np.zeros((K, N), np.float32)
        Wr[:kin, :nout] = rng.normal(0, scale * math.sqrt(2.0 / kin), (kin, nout))
        p[wo:wo + K * N] = Wr.ravel()
        br = np.zeros(N, np.float32)
        br[:nout] = rng.normal(0, 0.1, nout)
        p[bo:bo + N] = br
    return p


def random_norm(seed=0):
    rng = np.random.default_rng(seed + 1)
    nm = np.zeros(NORM, np.float32)
    nm[:N_IN] = rng.normal(0, 1, N_IN)
    nm[N_IN:2 * N_IN] = rng.uniform(0.5, 2.0, N_IN)
    nm[2 * N_IN:2 * N_IN + N_OUT] = rng.normal(0, 1, N_OUT)
    nm[2 * N_IN + N_OUT:2 * N_IN + 2 * N_OUT] = rng.uniform(0.5, 2.0, N_OUT)
    nm[2 * N_IN + 2 * N_OUT:3 * N_IN + 2 * N_OUT] = -1e30
    nm[3 * N_IN + 2 * N_OUT:] = 1e30
    return nm


# ------------------------------------------------------- features and labels --

def pair_frame(T_field: np.ndarray) -> np.ndarray:
    """Rows e1 (first edge), n x e1, n of a field triangle (vertices as rows):
    the frame of geo/calcoloSistemaRiferimento.m (e1 = (P2-P1)/|P2-P1|) with the
    normal of util/fileRead/readSpaceMesh.m:96-122."""
    P1, P2, P3 = T_field
    e1 = (P2 - P1) / np.linalg.norm(P2 - P1)
    n = np.cross(P1 - P3, P2 - P3)
    n = n / np.linalg.norm(n)
    return np.stack([e1, np.cross(n, e1), n])


def pair_features(Ts: np.ndarray, Tf: np.ndarray, lag: int, dt: float, cP: float, cS: float, h: float):
    """The 24 surrogate inputs of (test triangle Ts, field triangle Tf, lag)."""
    R = pair_frame(Tf)
    o = Tf.mean(0)
    x = np.zeros(N_IN)
    x[0:9] = ((Ts - o) @ R.T / h).ravel()
    x[9:18] = ((Tf - o) @ R.T / h).ravel()
    x[18], x[19], x[20] = dt * cP / h, dt * cS / h, lag
    d = np.linalg.norm(Ts.mean(0) - o)
    x[21] = (lag * dt * cP - d) / h
    x[22] = (lag * dt * cS - d) / h
    nS = np.cross(Ts[0] - Ts[2], Ts[1] - Ts[2])
    x[23] = (nS / np.linalg.norm(nS)) @ R[2]
    return x


def pair_labels(orc, S: int, F: int, lag: int, h: float):
    """P-front and S-front parts of E_lag[S, F] (no cull, no cleanup) from the
    SA oracle's per-node ris1..ris6 (SAc/BEMenerg_coreSA_calcMatrixSubBlock.m:
    101-111, 334-337: the P part is the ris1..ris4 terms, the S part the
    ris5, ris6 terms), rotated into the field triangle's frame and divided by
    A_S A_F / (4 pi rho h).  Returns (y (18,), E_total (3, 3) global)."""
    X, tri = orc.coords, orc.tri
    Ts, Tf = X[tri[S]], X[tri[F]]
    nF = np.cross(Tf[0] - Tf[2], Tf[1] - Tf[2])
    aF = np.linalg.norm(nF) / 2
    nF = nF / (2 * aF)
    aS = np.linalg.norm(np.cross(Ts[0] - Ts[2], Ts[1] - Ts[2])) / 2
    p = orc.param
    cP, cS, rho, dt = p.velP, p.velS, p.rho, p.Tfin / p.nT
    den = (cP * cS) ** 2
    EP = np.zeros((3, 3))
    ES = np.zeros((3, 3))
    for v, c in enumerate((1.0, -2.0, 1.0)):
        m = lag + 1 - v
        if m <= 0:
            continue
        t = m * dt
        for q in range(orc.nGH):
            x = orc.ghn[q] @ Ts
            w = aS * orc.ghw[q]
            ris, _ = orc.node_ris(x, Tf, nF, t)
            P = (ris[0] - ris[1]) / cP ** 2 + t * t * (-ris[2] + 3 * ris[3])
            Sp = ((cP ** 2 + cS ** 2) * ris[4] + (cP ** 2 - cS ** 2) * ris[5]) / den
            EP += c * w * 0.5 * P
            ES += c * w * 0.5 * Sp
    EP /= 4 * math.pi * rho
    ES /= 4 * math.pi * rho
    R = pair_frame(Tf)
    sc = aS * aF / (4 * math.pi * rho * h)
    y = np.concatenate([(R @ EP @ R.T).ravel(), (R @ ES @ R.T).ravel()]) / sc
    return y, EP + ES
