"""Gauss-Hammer outer rules of util/quadRules/GaussHammer_base.m with the map
of GaussHammerComposite(1, n) (util/quadRules/GaussHammerComposite.m:31-52):
barycentric columns reordered (b1, b3, b2) and weights doubled, so that the
weights sum to 1 (0.9999999999999 for the 13-digit 7-point table)."""
from __future__ import annotations

import numpy as np


def _base(n: int):
    if n == 3:
        v1, v2, v3 = 2 / 3, 1 / 6, 1 / 6
        g = np.array([[v1, v2, v3], [v3, v1, v2], [v2, v3, v1]])
        w = np.full(3, 1 / 6)
    elif n == 7:
        a1 = [0.1012865073235, 0.7974269853531, 0.1012865073235, 0.4701420641051, 0.4701420641051,
              0.0597158717898, 0.3333333333333]
        a2 = [0.1012865073235, 0.1012865073235, 0.7974269853531, 0.0597158717898, 0.4701420641051,
              0.4701420641051, 0.3333333333333]
        g = np.zeros((7, 3))
        for i in range(7):
            g[i, 0], g[i, 1] = a1[i], a2[i]
            g[i, 2] = 1 - g[i, 0] - g[i, 1]
        w = np.array([0.1259391805448 / 2] * 3 + [0.1323941527885 / 2] * 3 + [0.225 / 2])
    elif n == 12:
        g = np.zeros((12, 3))
        perm3 = [(0, 1, 2), (2, 0, 1), (1, 2, 0)]
        for s, c in enumerate((0.873821971016996, 0.501426509658179)):
            v = [c, (1 - c) / 2, (1 - c) / 2]
            for k in range(3):
                g[3 * s + k] = [v[perm3[k][q]] for q in range(3)]
        v = [0.636502499121399, 0.310352451033784, 0.0]
        v[2] = 1 - v[0] - v[1]
        perm6 = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
        for k in range(6):
            g[6 + k] = [v[perm6[k][q]] for q in range(3)]
        w1, w2 = 0.050844906370207 / 2, 0.116786275726379 / 2
        w3 = (0.5 - 3 * w1 - 3 * w2) / 6
        w = np.array([w1] * 3 + [w2] * 3 + [w3] * 6)
    elif n == 19:
        g = np.zeros((19, 3))
        w = np.zeros(19)
        g[0] = 1 / 3
        w[0] = 0.09713579628279610 / 2
        pa = [(0.02063496160252593, 0.48968251919873700), (0.1258208170141290, 0.4370895914929355),
              (0.6235929287619356, 0.1882035356190322), (0.91054097321109410, 0.04472951339445297)]
        pw = [0.09400410068141950 / 6, 0.2334826230143263 / 6, 0.2389432167816273 / 6, 0.07673302697609430 / 6]
        for s in range(4):
            b = 1 + 3 * s
            v1, v2 = pa[s]
            g[b:b + 3] = [[v1, v2, v2], [v2, v1, v2], [v2, v2, v1]]
            w[b:b + 3] = pw[s]
        v = [0.03683841205473626, 0.74119859878449800, 0.0]
        v[2] = 1 - v[0] - v[1]
        perm6 = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
        for k in range(6):
            g[13 + k] = [v[perm6[k][q]] for q in range(3)]
            w[13 + k] = 0.2597012362637364 / 12
    else:
        raise ValueError("Valori nodi esterni non validi")
    return g, w


def gauss_hammer(n: int):
    """(nodes n x 3, weights n) of GaussHammerComposite(1, n)."""
    g, w = _base(n)
    return g[:, [0, 2, 1]].copy(), 2.0 * w
