"""Pins of the oracle's direct post-processing u = G alpha - P g
(functions/BEMenerg_dirPP/BEMenerg_ppD_calc.m with CUDAkernels/kernelG.cu and
kernelP.cu), independent of its own formulas:

* rigid translation: u = c (constant in space and time) has zero traction,
  so the representation formula gives u(x) = -P c = c inside a closed
  surface and 0 outside once every front has passed (the static limit of the
  double layer, Somigliana's identity) -- a wrong sign, factor or dropped
  term of kernelP's nine entries breaks it;
* P blocks transform as tensors of the right parity under rotations and
  reflections of the mesh and the point;
* the G part is the first difference of the numeric single-layer M of
  kernelLayer.cu (pinned in test_oracle_ghc.py);
* causality: nothing arrives at a far point before r / cP."""
import numpy as np
import pytest


def _orc(mesh, **kw):
    from oracle import Oracle
    o = Oracle(mesh.coordinates, mesh.triangles, rho=1.0, mu=1.0, lam=2.0, Tfin=8.0, nT=32, **kw)
    o.set_pp_rule(16, 3)
    return o


def test_rigid_translation_inside_and_outside(sphere2):
    o = _orc(sphere2)
    N, Nv = sphere2.numberTriangles, sphere2.number_nodes
    rng = np.random.default_rng(0)
    for x, inside in (([0.1, -0.2, 0.15], 1), ([0.0, 0.0, 0.6], 1), ([0.3, 0.2, -0.5], 1), ([2.5, 0.3, -0.4], 0),
                      ([-1.6, 1.2, 0.4], 0)):
        c = rng.normal(size=3)
        g = np.zeros((3 * Nv, 32))
        for k in range(3):
            g[k::3] = c[k]
        for t in (8.0, 6.3):
            u = o.ppD_point(np.zeros((3 * N, 32)), g, np.array(x), t)
            assert np.abs(u - inside * c).max() < 1e-6 * np.abs(c).max(), (x, t, u)


def test_causality_and_nonpositive_time(sphere2):
    o = _orc(sphere2)
    N, Nv = sphere2.numberTriangles, sphere2.number_nodes
    rng = np.random.default_rng(1)
    d, g = rng.normal(size=(3 * N, 32)), rng.normal(size=(3 * Nv, 32))
    # the sphere lies within 1 of the origin: at x = (4, 0, 0) nothing arrives before (4 - 1) / cP
    assert np.all(o.ppD_point(d, g, np.array([4.0, 0, 0]), 0.95 * 3 / o.velP) == 0)
    assert np.any(o.ppD_point(d, g, np.array([4.0, 0, 0]), 6.0) != 0)
    assert np.all(o.ppD_point(d, g, np.array([0.1, 0, 0]), 0.0) == 0)
    assert np.all(o.ppD_point(d, g, np.array([0.1, 0, 0]), -1.0) == 0)


@pytest.mark.parametrize("proper", [True, False])
def test_p_blocks_are_tensors(sphere2, proper):
    rng = np.random.default_rng(2 if proper else 3)
    A = rng.normal(size=(3, 3))
    Q, R = np.linalg.qr(A)
    Q = Q * np.sign(np.diag(R))
    if (np.linalg.det(Q) > 0) != proper:
        Q[:, 0] = -Q[:, 0]
    sh = np.array([0.4, -0.1, 0.7])
    o = _orc(sphere2)
    from oracle import Oracle
    o2 = Oracle(sphere2.coordinates @ Q.T + sh, sphere2.triangles, rho=1.0, mu=1.0, lam=2.0, Tfin=8.0, nT=32)
    o2.set_pp_rule(16, 3)
    sign = 1.0 if proper else -1.0
    checked = 0
    for x in ([0.2, 0.1, -0.3], [1.4, -0.5, 0.2]):
        x = np.array(x)
        for s in rng.integers(sphere2.number_nodes, size=4):
            for t, y in ((2.0, 3), (3.1, 1), (5.0, 12)):
                P = o.pp_p_block(x, t, y, s)
                P2 = o2.pp_p_block(Q @ x + sh, t, y, s)
                nP = np.linalg.norm(P)
                if nP < 1e-12:
                    assert np.linalg.norm(P2) < 1e-10
                    continue
                assert np.linalg.norm(P2 - sign * Q @ P @ Q.T) < 1e-9 * nP
                checked += 1
    assert checked >= 10


def test_g_part_is_the_difference_of_the_numeric_layer(sphere2):
    o = _orc(sphere2)
    rng = np.random.default_rng(4)
    dt = o.dt
    for _ in range(20):
        x = rng.normal(size=3) * 0.8
        F = int(rng.integers(sphere2.numberTriangles))
        t = rng.uniform(0.3, 8.0)
        y = int(rng.integers(int(np.ceil(t / dt))))
        G = o.pp_g_block(x, t, y, F)
        M0 = o.pp_block_num(x, F, t - y * dt) if t - y * dt > 0 else np.zeros((3, 3))
        M1 = o.pp_block_num(x, F, t - (y + 1) * dt) if t - (y + 1) * dt > 0 else np.zeros((3, 3))
        ref = M0 - M1
        assert np.abs(G - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-14 / (4 * np.pi)) + 2e-14
