"""Pins of the oracle's numeric inner quadrature mode (the reference's GPU
path: functions/BEMenerg_dir/CUDAkernels/kernelV.cu off-diagonal,
BEMenerg_dir_calcSubBlockDiagV.m diagonal): quadrature rules by exactness,
the kernel by its trace identity and against the analytic (SA) oracle where
the integrand is smooth."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def orc(screen2):
    from oracle import Oracle
    o = Oracle(screen2.coordinates, screen2.triangles, Tfin=1.0, nT=16)
    o.set_ghc(64, 3, 256)
    return o


def _mean_monomial(a, b, c):
    # average of b1^a b2^b b3^c over a triangle
    return 2.0 * math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 2)


@pytest.mark.parametrize("nsub,nnod,deg", [(1, 3, 2), (4, 3, 2), (16, 7, 5), (64, 3, 2), (64, 1, 1)])
def test_composite_rule_exactness(orc, nsub, nnod, deg):
    x, w = orc.gh_composite(nsub, nnod)
    assert abs(w.sum() - 1) < 2e-13
    assert np.all(np.abs(x.sum(1) - 1) < 1e-14) and x.min() > 0
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                got = float((w * x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c).sum())
                assert abs(got - _mean_monomial(a, b, c)) < 2e-12, (a, b, c)


def test_composite_subtriangles_tile_the_triangle(orc):
    # 1-point composite rule: the 16 sub-triangle centroids, each weight 1/16,
    # all distinct and on the 1/12 lattice of a 4-way split
    x, w = orc.gh_composite(16, 1)
    assert np.allclose(w, 1 / 16)
    assert len({tuple(np.round(p, 12)) for p in x}) == 16
    assert np.allclose(np.round(x * 12) / 12, x, atol=1e-14)


@pytest.mark.parametrize("n", [16, 64, 256])
def test_collapsed_rule_exactness(orc, n):
    x, w = orc.doppio_gauss(n)
    n1 = int(round(math.sqrt(n)))
    assert abs(w.sum() - 1) < 1e-14
    # exact for b1^a b2^b b3^c with a + b <= 2 n1 - 1 and a + b + c <= 2 n1 - 2
    for a in range(3):
        for b in range(3):
            for c in range(2 * n1 - 2 - a - b + 1):
                got = float((w * x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c).sum())
                assert abs(got - _mean_monomial(a, b, c)) < 1e-13 * max(1, c), (a, b, c)


def test_collapsed_rule_resolves_the_vertex_singularity(orc):
    # mean of 1/|y - P| over the right triangle (0,0),(1,0),(0,1) with P at the
    # origin, the third barycentric vertex: 2 * sqrt(2) * asinh(1) / ... closed form
    x, w = orc.doppio_gauss(256)
    A, B, P = np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.zeros(3)
    y = x[:, 0:1] * A + x[:, 1:2] * B + x[:, 2:3] * P
    got = float((w / np.linalg.norm(y, axis=1)).sum())
    # integral of 1/r over the triangle = sqrt(2) asinh(1) ... in polar: int_0^{pi/2} 1/(cos+sin) = sqrt2*atanh(1/sqrt2)
    exact = 2.0 * math.sqrt(2.0) * math.atanh(1 / math.sqrt(2.0))  # (integral) / (area 1/2)
    assert abs(got - exact) < 1e-12 * exact


def test_nucleo_trace_identity(orc):
    rng = np.random.default_rng(0)
    cP, cS = orc.velP, orc.velS
    for _ in range(50):
        x = rng.normal(size=3)
        r = float(np.linalg.norm(x))
        for t, want in ((r / cP * 0.9, 0.0), (r / cP * 1.1 if r / cP * 1.1 < r / cS else r / cS * 0.99,
                                               1 / (r * cP ** 2)), (r / cS * 1.1, 1 / (r * cP ** 2) + 2 / (r * cS ** 2))):
            G = orc.nucleo(x, t)
            assert np.allclose(G, G.T, rtol=0, atol=1e-15 * np.abs(G).max() + 1e-300)
            assert abs(np.trace(G) - want) < 1e-12 * (1 + abs(want)), (t, np.trace(G), want)
    # before the P front the kernel is exactly zero
    assert not np.any(orc.nucleo(np.array([0.3, 0.2, 0.1]), 0.1))


def test_ghc_matches_analytic_where_the_integrand_is_smooth(orc, screen2):
    """Pairs and lags whose field triangle lies, for every outer node and all
    three levels, strictly between the S and the P front: the inner integrand
    is smooth there and the 64x3 composite rule reproduces the analytic block."""
    o = orc
    V = o.coords[o.tri]
    dt = o.dt
    X = (o.ghn[None, :, 0:1] * V[:, None, 0, :] + o.ghn[None, :, 1:2] * V[:, None, 1, :]) + \
        o.ghn[None, :, 2:3] * V[:, None, 2, :]
    found = 0
    worst = 0.0
    for S in range(0, o.N, 5):
        d = np.sqrt(((X[S][None, :, None, :] - V[:, None, :, :]) ** 2).sum(-1))  # F, q, v
        dmin = d.min(axis=(1, 2)) * 0.97  # margin: vertices bound the triangle's points from above only
        dmax = d.max(axis=(1, 2))
        for F in range(o.N):
            for lag in range(2, 15):
                if o.velS * (lag + 1) * dt < dmin[F] and dmax[F] < o.velP * (lag - 1) * dt:
                    E = o.ghc_pair_block(lag, S, F)
                    R, _ = o.pair_block(lag, S, F)
                    sc = np.abs(R).max()
                    if sc == 0:
                        continue
                    worst = max(worst, np.abs(E - R).max() / sc)
                    found += 1
    assert found > 50, found
    assert worst < 1e-6, worst


def test_ghc_diagonal_is_close_to_analytic(orc):
    """Self blocks: the collapsed rule against the analytic integral (the fronts
    cut the triangle, so agreement is at quadrature accuracy only)."""
    errs = []
    for S in (0, 17, 63):
        for lag in (0, 1, 2):
            E = orc.ghc_pair_block(lag, S, S)
            R, _ = orc.pair_block(lag, S, S)
            errs.append(np.abs(E - R).max() / np.abs(R).max())
    assert max(errs) < 5e-2, errs


def test_numeric_pp_block_matches_analytic_for_smooth_cases(orc, screen2):
    """Points and times for which the whole triangle lies strictly between the
    S and the P sphere: the numeric M of the GPU post-processing reproduces the
    analytic M of ppA."""
    o = orc
    o.set_pp_rule(64, 3)
    V = o.coords[o.tri]
    rng = np.random.default_rng(2)
    found, worst = 0, 0.0
    for _ in range(400):
        x = rng.uniform(-1.0, 1.0, 3)
        x[2] = rng.uniform(0.2, 1.0)
        j = int(rng.integers(o.N))
        d = np.linalg.norm(V[j] - x, axis=1)
        lo, hi = d.min() * 0.97, d.max()
        # S sphere inside the smallest distance, P sphere beyond the largest
        t_lo, t_hi = hi / o.velP, lo / o.velS
        if t_lo >= t_hi:
            continue
        tau = 0.5 * (t_lo + t_hi)
        A = o.pp_block(x, j, tau)
        B = o.pp_block_num(x, j, tau)
        sc = np.abs(A).max()
        if sc == 0:
            continue
        worst = max(worst, np.abs(A - B).max() / sc)
        found += 1
    assert found > 50 and worst < 2e-6, (found, worst)  # 64x3 quadrature error at ~h distance: 3.5e-7


def test_numeric_pp_point_is_causal(orc, screen2):
    o = orc
    o.set_pp_rule(64, 3)
    dens = np.random.default_rng(1).normal(size=(3 * o.N, o.nT))
    x = np.array([0.0, 0.0, 1.5])
    # the P front reaches the screen's nearest point (distance 1.5) at t = 1.5 / c_P
    assert not np.any(o.pp_point_num(dens, x, 0.8 * 1.5 / o.velP))
    assert np.any(o.pp_point_num(dens, x, 0.999))
