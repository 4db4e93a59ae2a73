"""Direct post-processing u = G alpha - P g on an MI355X, through the C ABI:
against the oracle (FP64 bar 1e-10 relative to the field's largest entry),
the rigid-translation identity, and the bar's closed-form solution after the
direct Dirichlet solve (the representation formula end to end)."""
import numpy as np
import pytest

from conftest import mesh_path

pytestmark = pytest.mark.gpu


def test_ppd_vs_oracle(sphere2):
    from elastodynamicebem_amd import api
    from oracle import Oracle
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=2.0, Tfin=8.0, nT=32)
    o = Oracle(sphere2.coordinates, sphere2.triangles, rho=1.0, mu=1.0, lam=2.0, Tfin=8.0, nT=32)
    o.set_pp_rule(16, 3)
    rng = np.random.default_rng(21)
    N, Nv = sphere2.numberTriangles, sphere2.number_nodes
    dens = rng.normal(size=(3 * N, 32))
    g = rng.normal(size=(3 * Nv, 32))
    pts = np.array([[0.1, -0.2, 0.15], [0.0, 0.0, 0.93], [1.3, 0.4, -0.2], [3.0, 0.0, 0.5]])
    times = np.array([8.0, 4.25, 2.0, 0.6, 0.0, -1.0])
    with api.Solver(sphere2, pb, 7) as s:
        s.set_pp_rule(16, 3)
        u = s.pp_eval_direct(pts, times, dens, g)
        ms = s.timings()["pp_direct_ms"]
        from elastodynamicebem_amd._lib import EbemError
        with pytest.raises(EbemError):
            s.pp_eval_direct(pts, np.array([8.5]), dens, g)
    ref = np.zeros_like(u)
    for p in range(pts.shape[0]):
        for k, t in enumerate(times):
            ref[:, p, k] = o.ppD_point(dens, g, pts[p], t)
    assert np.all(u[:, :, -2:] == 0)
    sc = np.abs(ref).max()
    err = np.abs(u - ref).max() / sc
    print(f"direct PP sphere2: {pts.shape[0]} points x {times.size} times in {ms:.1f} ms, max error {err:.2e}")
    assert err < 1e-10


def test_ppd_rigid_translation(sphere3):
    """Level-3 sphere: a constant vertex datum with zero density gives u = c
    inside and 0 outside once the fronts have passed (Somigliana's identity)."""
    from elastodynamicebem_amd import api
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=6.0, nT=48)
    Nv, N = sphere3.number_nodes, sphere3.numberTriangles
    c = np.array([0.3, -1.2, 0.7])
    g = np.zeros((3 * Nv, 48))
    for k in range(3):
        g[k::3] = c[k]
    pts = np.array([[0.0, 0.0, 0.0], [0.2, 0.5, -0.3], [0.0, 0.0, 2.5], [-1.8, 0.9, 0.0]])
    inside = np.array([1, 1, 0, 0])
    with api.Solver(sphere3, pb, 7) as s:
        s.set_pp_rule(64, 3)
        u = s.pp_eval_direct(pts, np.array([6.0]), np.zeros((3 * N, 48)), g)[:, :, 0]
    err = np.abs(u - np.outer(c, inside)).max() / np.abs(c).max()
    print(f"rigid translation, level-3 sphere: max error {err:.2e}")
    assert err < 1e-6


def _bar_u(h, cP, Tfin, z, t):
    khat = int(np.ceil(cP * Tfin / (2 * h))) - 1
    k = np.arange(khat + 1)[:, None]
    a = cP * t - 2 * h * k - (h - z)
    b = cP * t - 2 * h * (k + 1) + (h - z)
    return np.sum((-1.0) ** k * (a * (a > 0) - b * (b > 0)), axis=0)


def test_ppd_bar_after_dd_solve():
    """barH1 level 2 (768 triangles, 96 steps): density from the direct
    Dirichlet driver, then u inside the bar from the direct post-processing
    against the closed form u_z(z, t) (u_x = u_y = 0), to the discretisation
    error."""
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd.mesh import read_mesh
    bar = read_mesh(mesh_path("barH1_2"))
    pb = api.PbParam(rho=0.25, mu=0.5, lambda_=0.0, Tfin=6.0, nT=96, domainType="barH1", lev=2)
    pts = np.array([[0.0, 0.0, 0.25], [0.1, -0.2, 0.5], [-0.2, 0.15, 0.8]])
    times = np.array([0.8, 1.5, 2.6, 4.0])
    with api.Solver(bar, pb, 7) as s:
        s.set_inner_rule(16, 3, 64)
        s.set_pp_rule(16, 3)
        s.assemble(api.PREC_GHC)
        s.assemble_k()
        s.rhs_dirichlet(api.DATUM_REF_BAR_H1)
        s.mot_solve(None)
        u = s.pp_eval_direct(pts, times)  # device density and datum
    ex = np.zeros_like(u)
    for k, t in enumerate(times):
        ex[2, :, k] = _bar_u(1.0, 2.0, 6.0, pts[:, 2], t)
    rel = np.linalg.norm(u - ex) / np.linalg.norm(ex)
    print(f"bar interior field after the DD solve: rel L2 error {rel:.3e} (max |u_x|, |u_y|: "
          f"{np.abs(u[:2]).max():.2e}, max |u_z| {np.abs(ex[2]).max():.2f})")
    assert rel < 0.03


def test_ppd_per_vertex_path_agrees(sphere2, monkeypatch):
    """The per-vertex double layer (EBEM_PPD_BY_FIELD=0: k_ppd_P, the
    reference's accumulation order and kernel expressions) and the default by
    field triangle and level (k_ppf + k_ppf_gather, reciprocal forms of the
    kernel) agree to the FP64 bar.  The kernel's terms are much larger than
    their sum, so an ulp in a node's value moves the field by ~1e-11 of its
    size: that is the reference algorithm's own rounding noise, which the
    bit-identical per-vertex kernel shares with the oracle."""
    from elastodynamicebem_amd import api
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=2.0, Tfin=8.0, nT=32)
    rng = np.random.default_rng(22)
    N, Nv = sphere2.numberTriangles, sphere2.number_nodes
    dens = rng.normal(size=(3 * N, 32))
    g = rng.normal(size=(3 * Nv, 32))
    pts = np.array([[0.1, -0.2, 0.15], [0.0, 0.0, 0.93], [1.3, 0.4, -0.2]])
    times = np.array([8.0, 5.1, 2.0, 0.6])
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("EBEM_PPD_BY_FIELD", mode)
        with api.Solver(sphere2, pb, 7) as s:
            s.set_pp_rule(16, 3)
            out[mode] = s.pp_eval_direct(pts, times, dens, g)
    err = np.abs(out["0"] - out["1"]).max() / np.abs(out["0"]).max()
    print(f"direct PP per vertex vs by field triangle and level: {err:.2e}")
    assert err < 1e-10
