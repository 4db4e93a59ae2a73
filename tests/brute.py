"""Independent numerical integrator for the single-layer kernel integrals.

For an outer point x, a flat field triangle T and a time t it integrates,
over T, the six kernels the analytic routines of the reference produce
(SAc/BEMenerg_coreSA_calcMatrixSubBlock.m:34-39):

  ris1 = I * int_{T, cS t < r < cP t} 1/r          ris2 = int_{same} d d^T / r^3
  ris3 = I * int_{same} 1/r^3                      ris4 = int_{same} d d^T / r^5
  ris5 = I * int_{T, r < cS t} 1/r                 ris6 = int_{same} d d^T / r^3

with d = y - x and r = |d|.  The method shares nothing with the analytic
formulas: polar coordinates about the projection O of x on the plane of T,
the triangle split into the three fans (O, V_k, V_k+1) with signed sweep
angles, composite Gauss-Legendre in the angle (split at every angle where a
wave front meets an edge line) and in the radius (split at the fronts and at
geometric multiples of the plane distance).  Converges to ~1e-13 relative.
"""
from __future__ import annotations

import numpy as np

_GX, _GW = np.polynomial.legendre.leggauss(40)


def _gl(a, b):
    x = 0.5 * (b - a) * _GX + 0.5 * (b + a)
    return x, 0.5 * (b - a) * _GW


def _radial_nodes(a, b, zeta):
    """Composite GL nodes on [a, b], refined near rho ~ zeta."""
    cuts = [a, b]
    if zeta > 0:
        s = zeta
        while s < b:
            if s > a:
                cuts.append(s)
            s *= 4.0
    cuts = sorted(set(cuts))
    xs, ws = [], []
    for lo, hi in zip(cuts[:-1], cuts[1:]):
        if hi > lo:
            x, w = _gl(lo, hi)
            xs.append(x)
            ws.append(w)
    return np.concatenate(xs), np.concatenate(ws)


def kernel_integrals(x, T, t, cP, cS):
    """Return ris (6, 3, 3) in the global frame."""
    x = np.asarray(x, float)
    T = np.asarray(T, float)
    n = np.cross(T[1] - T[0], T[2] - T[0])
    n /= np.linalg.norm(n)
    e1 = (T[1] - T[0]) / np.linalg.norm(T[1] - T[0])
    e2 = np.cross(n, e1)
    s = np.dot(x - T[0], n)             # signed height of x above the plane
    zeta = abs(s)
    O = x - s * n
    P2 = np.array([[np.dot(v - O, e1), np.dot(v - O, e2)] for v in T])
    orient = np.sign((P2[1, 0] - P2[0, 0]) * (P2[2, 1] - P2[0, 1]) - (P2[1, 1] - P2[0, 1]) * (P2[2, 0] - P2[0, 0]))
    rP, rS = cP * t, cS * t
    rhoP = np.sqrt(rP * rP - zeta * zeta) if rP > zeta else 0.0
    rhoS = np.sqrt(rS * rS - zeta * zeta) if rS > zeta else 0.0
    acc = np.zeros((6, 3, 3))
    for k in range(3):
        A, B = P2[k], P2[(k + 1) % 3]
        cr = A[0] * B[1] - A[1] * B[0]
        if abs(cr) < 1e-300:
            continue
        thA = np.arctan2(A[1], A[0])
        dth = np.arctan2(cr, A[0] * B[0] + A[1] * B[1])
        u = (B - A) / np.linalg.norm(B - A)
        foot = A - np.dot(A, u) * u
        h = np.linalg.norm(foot)
        thn = np.arctan2(foot[1], foot[0])
        lo, hi = (thA, thA + dth) if dth > 0 else (thA + dth, thA)
        cuts = [lo, hi]
        for R in (rhoP, rhoS):
            if R > h:
                ph = np.arccos(h / R)
                for c in (thn + ph, thn - ph):
                    c = c + 2 * np.pi * np.round((0.5 * (lo + hi) - c) / (2 * np.pi))
                    if lo < c < hi:
                        cuts.append(c)
        cuts = sorted(cuts)
        thl, wtl = [], []
        for a, b in zip(cuts[:-1], cuts[1:]):
            # geometric grading toward both ends of every piece: the edge
            # distance h/cos varies on the scale h/R next to a front crossing
            fr = [0.0, 1e-6, 1e-5, 1e-4, 1e-3, 1e-2, 0.05, 0.2, 0.5]
            sub = sorted(set([a + f * (b - a) for f in fr] + [b - f * (b - a) for f in fr]))
            for a2, b2 in zip(sub[:-1], sub[1:]):
                x_, w_ = _gl(a2, b2)
                thl.append(x_)
                wtl.append(w_)
        if thl:
            for th, wt in zip(np.concatenate(thl), np.concatenate(wtl)):
                redge = h / np.cos(th - thn)
                dirv = np.cos(th) * e1 + np.sin(th) * e2
                for (r0, r1, kern) in ((rhoS, min(rhoP, redge), 0), (0.0, min(rhoS, redge), 1)):
                    if kern == 0 and rP <= zeta:
                        continue
                    if kern == 1 and rS <= zeta:
                        continue
                    if r1 <= r0:
                        continue
                    rho, wr = _radial_nodes(r0, r1, zeta)
                    d = rho[:, None] * dirv[None, :] - s * n[None, :]
                    r = np.sqrt(rho * rho + zeta * zeta)
                    w = wr * rho * wt * np.sign(dth)
                    dd = d[:, :, None] * d[:, None, :]
                    if kern == 0:
                        acc[0] += np.eye(3) * np.sum(w / r)
                        acc[1] += np.einsum("i,ijk->jk", w / r ** 3, dd)
                        acc[2] += np.eye(3) * np.sum(w / r ** 3)
                        acc[3] += np.einsum("i,ijk->jk", w / r ** 5, dd)
                    else:
                        acc[4] += np.eye(3) * np.sum(w / r)
                        acc[5] += np.einsum("i,ijk->jk", w / r ** 3, dd)
    return acc * orient


def combine(ris, t, cP, cS, rho=1.0):
    """Single-layer kernel combination (calcMatrixSubBlock.m:101-104) of one level."""
    vP2, vS2 = cP * cP, cS * cS
    out = ((ris[0] - ris[1]) / vP2 + t * t * (-ris[2] + 3 * ris[3])
           + ((vP2 + vS2) * ris[4] + (vP2 - vS2) * ris[5]) / (cP * cS) ** 2) / 2
    return out / (4 * np.pi * rho)
