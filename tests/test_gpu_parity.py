"""Parity of the HIP library (through its C ABI) with the oracle, on an MI355X.

Bar: FP64, Frobenius-relative 1e-10 per E_lag (absolute 1e-10 * ||E_0|| for
lags whose norm is below 1e-6 ||E_0||), 1e-10 on the density.  Where the
reference algorithm itself is less reproducible than that -- its closed forms
are ill-conditioned for thin child fans, so a 1-ulp change of the mesh
coordinates moves e.g. E_30 of the level-3 sphere by 2.5e-9 -- the bar is 3x
that measured sensitivity (`noise`, computed here with the oracle on
coordinates nudged by one ulp)."""
import numpy as np
import pytest

from conftest import frob_rel

pytestmark = pytest.mark.gpu


def _solver(mesh, Tfin, nT, nGH=7):
    from elastodynamicebem_amd import api
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=Tfin, nT=nT)
    return api.Solver(mesh, pb, nGH)


def _oracle(mesh, Tfin, nT, nGH=7):
    from oracle import Oracle
    return Oracle(mesh.coordinates, mesh.triangles, Tfin=Tfin, nT=nT, nGH=nGH)


def _ulp_oracle(mesh, Tfin, nT, nGH=7, seed=0):
    """Oracle on coordinates moved by one ulp (random direction per entry)."""
    from oracle import Oracle
    rng = np.random.default_rng(seed)
    X = mesh.coordinates
    X2 = np.where(rng.random(X.shape) < 0.5, np.nextafter(X, np.inf), np.nextafter(X, -np.inf))
    return Oracle(X2, mesh.triangles, Tfin=Tfin, nT=nT, nGH=nGH)


def _check_lags(s, o, lags, rows=None, o_ulp=None):
    E0n = None
    worst = 0.0
    for l in lags:
        G = s.get_block(l)
        R, _ = o.matrix_block(l, rows=rows)
        R2 = o_ulp.matrix_block(l, rows=rows)[0] if o_ulp is not None else None
        if rows is not None:
            G = G[3 * rows[0]:3 * rows[1]]
            R = R[3 * rows[0]:3 * rows[1]]
            if R2 is not None:
                R2 = R2[3 * rows[0]:3 * rows[1]]
        if E0n is None:
            E0n = np.linalg.norm(R)
        nr = np.linalg.norm(R)
        if nr >= 1e-6 * E0n:
            e = frob_rel(G, R)
            noise = frob_rel(R2, R) if R2 is not None else 0.0
        else:
            e = np.linalg.norm(G - R) / E0n
            noise = np.linalg.norm(R2 - R) / E0n if R2 is not None else 0.0
        assert e < max(1e-10, 3 * noise), (l, e, noise)
        # identical sparsity up to cleanup-threshold ties
        assert np.count_nonzero((G != 0) != (R != 0)) <= 1e-4 * G.size, l
        worst = max(worst, e)
    return worst


def test_node_integral_vs_oracle(sphere2):
    s = _solver(sphere2, 4.0, 32)
    o = _oracle(sphere2, 4.0, 32)
    rng = np.random.default_rng(3)
    errs = []
    for k in range(200):
        j = int(rng.integers(o.N))
        x = o.center[int(rng.integers(o.N))] + rng.normal(size=3) * 0.02 * (k % 3)
        t = rng.uniform(0.01, 1.5)
        ref = o.pp_block(x, j, t)
        g = s.node_integral(x, j, t)
        g[np.abs(g) < 1e-14] = 0
        sc = np.abs(ref).max()
        if sc == 0:
            assert np.abs(g).max() == 0
            continue
        errs.append(np.abs(g - ref).max() / sc)
    errs = np.array(errs)
    assert np.median(errs) < 1e-12 and errs.max() < 1e-8


def test_cfg0_blocks_rhs_mot(screen2):
    """Plain method, CPU-runnable case: 8x8 screen, 128 triangles, 16 steps on [0,1]."""
    s = _solver(screen2, 1.0, 16)
    o = _oracle(screen2, 1.0, 16)
    L = s.num_blocks()
    assert L == o.num_blocks() == 16
    s.assemble()
    info = s.store_info()
    # culled pair.lags of the reference cull (BASELINE workload table: 2.055e5)
    assert abs(info["pair_lags"] - 2.055e5) / 2.055e5 < 5e-3
    _check_lags(s, o, range(L))
    beta_g = s.rhs(0)
    beta_o = o.rhs(0)
    assert frob_rel(beta_g, beta_o) < 1e-14
    E = np.stack([o.matrix_block(l)[0] for l in range(L)])
    a_o = o.mot(E, beta_o)
    a_g = s.mot_solve(None)
    assert frob_rel(a_g, a_o) < 1e-10


def test_sphere2_all_lags_and_mot(sphere2):
    """Level-2 icosphere (320 triangles), t in [0,4], 32 steps, datum n t^2 e^{-4t}."""
    s = _solver(sphere2, 4.0, 32)
    o = _oracle(sphere2, 4.0, 32)
    L = s.num_blocks()
    assert L == o.num_blocks()
    s.assemble()
    _check_lags(s, o, range(L))
    beta_o = o.rhs(1)
    assert frob_rel(s.rhs(1), beta_o) < 1e-14
    E = np.stack([o.matrix_block(l)[0] for l in range(L)])
    a_o = o.mot(E, beta_o)
    a_g = s.mot_solve(None)
    assert frob_rel(a_g, a_o) < 1e-10


@pytest.mark.parametrize("nGH", [3, 12, 19])
def test_other_outer_rules(screen2, nGH):
    s = _solver(screen2, 1.0, 16, nGH)
    o = _oracle(screen2, 1.0, 16, nGH)
    s.assemble()
    _check_lags(s, o, [0, 3, 9])


def test_sphere3_full_size_sampled_rows(sphere3):
    """Full-size sphere (1280 triangles, 128 steps on [0,4], L=66): sampled test
    rows of E_lag against the oracle, and the store size against the cull."""
    s = _solver(sphere3, 4.0, 128)
    o = _oracle(sphere3, 4.0, 128)
    assert s.num_blocks() == 66
    s.assemble()
    info = s.store_info()
    assert abs(info["pair_lags"] - 5.614e7) / 5.614e7 < 5e-3
    o2 = _ulp_oracle(sphere3, 4.0, 128)
    for rows in ((0, 6), (640, 646), (1274, 1280)):
        _check_lags(s, o, [0, 1, 7, 30, 64, 65], rows=rows, o_ulp=o2)


def test_pp_vs_oracle(screen2, sphere2):
    """Analytic post-processing u(x,t) (ppA semantics) against the oracle, on grid
    times and off-grid times, for points near and far from the boundary."""
    rng = np.random.default_rng(5)
    for mesh, Tfin, nT in ((screen2, 1.0, 16), (sphere2, 4.0, 32)):
        s = _solver(mesh, Tfin, nT)
        o = _oracle(mesh, Tfin, nT)
        dens = rng.normal(size=(3 * o.N, nT))
        pts = np.concatenate([o.center[rng.integers(o.N, size=3)] + rng.normal(size=(3, 3)) * 0.3,
                              rng.normal(size=(2, 3)) * 1.5])
        dt = Tfin / nT
        times = np.concatenate([dt * np.arange(1, nT + 1), rng.uniform(0, Tfin, size=5), [0.0, -1.0]])
        u = s.pp_eval(pts, times, dens)  # (3, npts, nt)
        ref = np.zeros_like(u)
        ref2 = np.zeros_like(u)
        pts2 = np.nextafter(pts, np.inf)
        for p in range(pts.shape[0]):
            for k, tk in enumerate(times):
                ref[:, p, k] = o.pp_point(dens, pts[p], tk)
                ref2[:, p, k] = o.pp_point(dens, pts2[p], tk)
        assert np.all(u[:, :, -2:] == 0)
        noise = frob_rel(ref2, ref)
        assert frob_rel(u, ref) < max(1e-10, 3 * noise), (frob_rel(u, ref), noise)


def test_pp_point_order_does_not_matter(sphere2):
    """The post-processing takes the points in its own (Morton) order and
    pairs neighbours in the convolution; each point's sum is its own, so any
    order of the caller's points gives the same field bit for bit, written
    back to the caller's columns (an odd count leaves one point alone)."""
    rng = np.random.default_rng(11)
    Tfin, nT = 4.0, 32
    s = _solver(sphere2, Tfin, nT)
    dens = rng.normal(size=(3 * sphere2.numberTriangles, nT))
    v = rng.normal(size=(37, 3))
    pts = v / np.linalg.norm(v, axis=1)[:, None] * rng.uniform(1.1, 2.0, size=(37, 1))
    times = Tfin / nT * np.arange(1, nT + 1)
    u = s.pp_eval(pts, times, dens)
    perm = rng.permutation(pts.shape[0])
    up = s.pp_eval(pts[perm], times, dens)
    assert np.array_equal(up, u[:, perm, :])
    assert np.abs(u).max() > 0
    s.close()


def test_pp_after_mot_uses_device_density(screen2):
    s = _solver(screen2, 1.0, 16)
    s.assemble()
    s.rhs(0)
    a = s.mot_solve(None)
    x = np.array([[0.0, 0.0, 0.3], [0.2, -0.1, 0.5]])
    t = np.array([0.5, 1.0])
    u1 = s.pp_eval(x, t)          # device density of the MOT
    u2 = s.pp_eval(x, t, a)       # same density from the host
    assert np.array_equal(u1, u2)


def test_e0_unpivoted_and_pivoted_factorisations_agree(sphere2, monkeypatch):
    """E_0 is inverted by LU without pivoting when that passes the residual
    check, else with partial pivoting; both give the same density."""
    from elastodynamicebem_amd import api
    s = _solver(sphere2, 4.0, 32)
    s.assemble()
    s.rhs(1)
    a0 = s.mot_solve(None)
    monkeypatch.setenv("EBEM_E0_PIVOT", "1")
    a1 = s.mot_solve(None)
    assert frob_rel(a0, a1) < 1e-12
    s.close()


def test_pp_conv_step_runs_and_gaps_agree(sphere2):
    """The same times asked for as one run of steps and as a set with gaps
    (other time groups, other convolution passes) give the same field."""
    rng = np.random.default_rng(6)
    Tfin, nT = 4.0, 32
    s = _solver(sphere2, Tfin, nT)
    dens = rng.normal(size=(3 * sphere2.numberTriangles, nT))
    pts = np.concatenate([rng.normal(size=(3, 3)) * 0.4, rng.normal(size=(2, 3)) * 1.5])
    dt = Tfin / nT
    steps = np.array([2, 5, 6, 9, 20, 31])
    u_all = s.pp_eval(pts, dt * np.arange(1, nT + 1), dens)
    u_gap = s.pp_eval(pts, dt * steps, dens)
    err = frob_rel(u_gap, u_all[:, :, steps - 1])
    print(f"post-processing, consecutive steps vs steps with gaps: {err:.2e}")
    assert err < 1e-12
