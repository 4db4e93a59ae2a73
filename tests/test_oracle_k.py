"""Pins of the oracle's double-layer operator K (the reference's direct
Dirichlet driver: functions/BEMenerg_dir/CUDAkernels/kernelK.cu for the
regular part, BEMenerg_dir_calcSubBlockSingK.m + calcKIntgSing.m for a test
triangle against its own vertices).

What fixes the restatement here, independently of its own formulas:
* the P1 basis by its defining values phi_k(V_l) = delta_kl;
* every block transforms as a tensor of the right parity: under a proper
  rotation Q of the mesh K -> Q K Q^T, under a reflection (vertex order kept,
  so the normals become -Q n) K -> -Q K Q^T.  A wrong index, a transposed
  factor or a sign slip in one of the nine hand-written components of the
  reference's kernels breaks this;
* past the S-wave front every part of the kernel is linear in t (its time
  antiderivative is affine), so the third time difference [-1, 3, -3, 1]
  cancels: blocks that the reference's skip tests drop are zero when they
  are integrated anyway.  A dropped or mistyped power of (t - r/c) breaks it.
The driver-level known answer (the 1-D bar's closed form) is checked on the
device."""
import numpy as np
import pytest


def _orc(coords, tri, **kw):
    from oracle import Oracle
    o = Oracle(coords, tri, rho=1.0, mu=1.0, lam=2.0, Tfin=4.0, nT=32, **kw)
    o.set_ghc(16, 3, 64)
    return o


def _rot(rng, proper=True):
    A = rng.normal(size=(3, 3))
    Q, R = np.linalg.qr(A)
    Q = Q * np.sign(np.diag(R))
    if (np.linalg.det(Q) > 0) != proper:
        Q[:, 0] = -Q[:, 0]
    return Q


def _pairs(o, rng, n):
    """(S, s) pairs: n with s a vertex of S (singular part) and n with s a
    vertex of a neighbour but not of S, plus n at random."""
    tri = o.tri
    out = []
    for S in rng.choice(o.N, n, replace=False):
        out.append((int(S), int(tri[S][rng.integers(3)])))
        nb = [F for F in range(o.N) if F != S and len(set(tri[F]) & set(tri[S])) > 0]
        F = nb[rng.integers(len(nb))]
        others = [v for v in tri[F] if v not in tri[S]]
        if others:
            out.append((int(S), int(others[0])))
        out.append((int(S), int(rng.integers(o.coords.shape[0]))))
    return out


def test_basis_values(sphere2):
    o = _orc(sphere2.coordinates, sphere2.triangles)
    rng = np.random.default_rng(0)
    for _ in range(50):
        T = rng.normal(size=(3, 3))
        mat, vec = o.k_basis(T)
        phi = vec[:, None] + mat @ T.T          # phi[k, l] = phi_k(V_l)
        assert np.abs(phi - np.eye(3)).max() < 1e-12
        n = np.cross(T[1] - T[0], T[2] - T[0])
        assert np.abs(mat @ n).max() < 1e-12 * np.abs(mat).max() * np.linalg.norm(n)
        assert abs(vec.sum() - 1) < 1e-12 and np.abs(mat.sum(0)).max() < 1e-12 * np.abs(mat).max()


@pytest.mark.parametrize("proper", [True, False])
def test_blocks_are_tensors(sphere2, proper):
    rng = np.random.default_rng(1 if proper else 2)
    X = sphere2.coordinates
    Q = _rot(rng, proper)
    o = _orc(X, sphere2.triangles)
    o2 = _orc(X @ Q.T + np.array([0.3, -0.7, 1.1]), sphere2.triangles)
    sign = 1.0 if proper else -1.0
    checked = 0
    for S, s in _pairs(o, rng, 6):
        lags = (0, 1, 2, 5, 11)
        K = [o.k_block(l, S, s, prune=False) for l in lags]
        K2 = [o2.k_block(l, S, s, prune=False) for l in lags]
        # the singular part cancels large terms: compare on the scale of the
        # pair's largest block (late blocks are round-off around zero)
        peak = max(np.linalg.norm(k) for k in K)
        for l, k, k2 in zip(lags, K, K2):
            if np.linalg.norm(k) < 1e-8 * peak:
                assert np.linalg.norm(k2) < 1e-8 * peak
                continue
            e = np.linalg.norm(k2 - sign * Q @ k @ Q.T) / peak
            assert e < 1e-10, (S, s, l, e)
            checked += 1
    assert checked >= 40


def test_skipped_blocks_are_zero(sphere2):
    """Blocks past both fronts (isBlockNull's first test, and the singular
    blocks' cut-off) vanish when integrated; so do blocks before the P front
    once all four time levels are positive (lag >= 3)."""
    o = _orc(sphere2.coordinates, sphere2.triangles)
    rng = np.random.default_rng(3)
    dt, cP, cS = o.dt, o.velP, o.velS
    maxLen = o.maxL.max()
    X = o.coords
    late = early = 0
    for S, s in _pairs(o, rng, 8):
        d = np.linalg.norm(X[o.tri[S]] - X[s], axis=1)
        peak = max(np.linalg.norm(o.k_block(l, S, s, prune=False)) for l in range(0, 8))
        assert peak > 0
        first_late = int(np.floor((d.max() + maxLen) / (cS * dt))) + 3
        if s in o.tri[S]:
            first_late = max(first_late, int(np.ceil(o.maxL[S] / (cS * dt))) + 2)
        for lag in (first_late, first_late + 4):
            if lag >= o.nT:
                continue
            assert np.abs(o.k_block(lag, S, s, prune=False)).max() < 1e-10 * peak, (S, s, lag)
            late += 1
        for lag in range(3, o.nT):
            if (lag + 1) * cP * dt < d.min() - 2 * maxLen:
                assert np.abs(o.k_block(lag, S, s, prune=False)).max() < 1e-10 * peak, (S, s, lag)
                early += 1
    assert late >= 10 and early >= 1


def test_pruned_matrix_matches_blocks(screen2):
    """Dense K_lag rows and single blocks agree; once every pair is past its
    fronts the sum of K_lag over all lags vanishes (it telescopes to a second
    difference of the affine late-time antiderivative).  With lambda = 0 the
    kernelK.cu Heaviside typo ((r - r/cP) > 0 for (t - r/cP) > 0 in the
    lambda-mu term) has no weight, and the skip tests drop zeros only."""
    from oracle import Oracle
    o = Oracle(screen2.coordinates, screen2.triangles, rho=1.0, mu=1.0, lam=0.0, Tfin=4.0, nT=16)
    o.set_ghc(16, 3, 64)
    LK = o.k_num_blocks()
    assert LK == min(o.num_blocks() + 1, 16) < 16
    rows = (40, 43)
    Ks = [o.k_matrix(l, rows) for l in range(16)]
    rng = np.random.default_rng(4)
    for S in range(*rows):
        for s in rng.integers(o.coords.shape[0], size=4):
            for l in (0, 2, 7):
                assert np.array_equal(Ks[l][3 * S:3 * S + 3, 3 * s:3 * s + 3], o.k_block(l, S, s))
    tot = sum(Ks)[3 * rows[0]:3 * rows[1]]
    assert np.abs(tot).max() < 1e-10 * max(np.abs(K).max() for K in Ks)
    # the reference keeps numBlocksK = min(L + 1, nT) lags: the rest are zero
    assert all(np.abs(K).max() == 0 for K in Ks[LK:])


def test_typo_weight_before_the_p_front(screen2):
    """With lambda != 0 the typo gives the regular kernel a value before the
    P wave arrives; it cancels in the third difference from lag 3 on, so only
    lags 0..2 carry it, and there the reference's early skip test drops it
    for pairs beyond the front (kept as the reference does)."""
    from oracle import Oracle
    o = Oracle(screen2.coordinates, screen2.triangles, rho=1.0, mu=1.0, lam=2.0, Tfin=4.0, nT=16)
    o.set_ghc(16, 3, 64)
    o0 = Oracle(screen2.coordinates, screen2.triangles, rho=1.0, mu=1.0, lam=0.0, Tfin=4.0, nT=16)
    o0.set_ghc(16, 3, 64)
    S, s = 40, 4     # 0.90 from the nearest vertex: beyond the P front at lag 0
    assert np.abs(o.k_block(0, S, s)).max() == 0
    assert np.abs(o.k_block(0, S, s, prune=False)).max() > 1e-7
    assert np.abs(o0.k_block(0, S, s, prune=False)).max() == 0


def test_beta_k_is_the_causal_convolution(screen2):
    from oracle import Oracle
    o = Oracle(screen2.coordinates, screen2.triangles, Tfin=1.0, nT=8)
    o.set_ghc(4, 3, 16)
    rng = np.random.default_rng(5)
    g = rng.normal(size=(3 * o.coords.shape[0], 8))
    rows = (0, 3)
    b = o.beta_k(g, rows)
    K = [o.k_matrix(l, rows) for l in range(o.k_num_blocks())]
    n = 5
    ref = sum(K[n - j] @ g[:, j] for j in range(n + 1) if n - j < len(K))
    assert np.allclose(b[:, n], ref, rtol=0, atol=1e-14 * np.abs(ref).max())
    # causality: a datum that starts at step 3 leaves beta_0..2 untouched
    g2 = g.copy()
    g2[:, :3] = 0
    assert np.all(o.beta_k(g2, rows)[:, :3] == 0)
