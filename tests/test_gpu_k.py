"""Double-layer operator K and the direct Dirichlet driver on an MI355X,
through the C ABI, against the oracle and the 1-D bar's closed form.

Bar: FP64, per lag the Frobenius error of the sampled test rows of K_lag
relative to the largest of those rows over all lags below 1e-10 (late blocks
are small third differences whose round-off is set by the early ones); beta_K
and the combined right-hand side 1e-12 relative."""
import numpy as np
import pytest

from conftest import frob_rel, mesh_path

pytestmark = pytest.mark.gpu


def _pb(Tfin, nT, rho=1.0, mu=1.0, lam=2.0, domain=None):
    from elastodynamicebem_amd import api
    pb = api.PbParam(rho=rho, mu=mu, lambda_=lam, Tfin=Tfin, nT=nT)
    if domain:
        pb.domainType = domain
    return pb


def _check_k_rows(s, o, rows, LK):
    Ks = [s.get_k_block(l, rows) for l in range(o.nT)]
    Rs = [o.k_matrix(l, rows)[3 * rows[0]:3 * rows[1]] for l in range(LK)]
    peak = max(np.linalg.norm(R) for R in Rs)
    assert peak > 0
    worst = 0.0
    for l in range(LK):
        e = np.linalg.norm(Ks[l] - Rs[l]) / peak
        worst = max(worst, e)
        assert e < 1e-10, (rows, l, e)
        # identical sparsity up to clean-up-threshold ties
        assert np.count_nonzero((Ks[l] != 0) != (Rs[l] != 0)) <= 1e-3 * Ks[l].size + 2, l
    for l in range(LK, o.nT):
        assert np.all(Ks[l] == 0)
    return worst


@pytest.mark.parametrize("nGH,nsub,nsing,ndiag", [(7, 16, 3, 64), (19, 64, 3, 256)])
def test_k_rows_vs_oracle_sphere(sphere2, nGH, nsub, nsing, ndiag):
    """Level-2 sphere, lambda = 2 (the kernelK.cu typo term has weight),
    t in [0, 4], 32 steps; the reference's default FN rule 19 64 3 256 too."""
    from elastodynamicebem_amd import api
    from oracle import Oracle
    pb = _pb(4.0, 32)
    o = Oracle(sphere2.coordinates, sphere2.triangles, rho=1.0, mu=1.0, lam=2.0, Tfin=4.0, nT=32, nGH=nGH)
    o.set_ghc(nsub, nsing, ndiag)
    with api.Solver(sphere2, pb, nGH) as s:
        s.set_inner_rule(nsub, nsing, ndiag)
        LK = s.k_num_blocks()
        assert LK == o.k_num_blocks()
        s.assemble_k()
        info = s.k_store_info()
        assert 0 < info["pairs"] <= o.N * sphere2.number_nodes and info["slots"] >= info["pairs"]
        rows = [(0, 3), (157, 160)] if nGH == 7 else [(40, 42)]
        worst = max(_check_k_rows(s, o, r, LK) for r in rows)
        print(f"K sphere2 FN {nGH} {nsub} {nsing} {ndiag}: LK {LK}, {info['slots']} slots, "
              f"{s.timings()['k_assembly_ms']:.1f} ms, worst lag error {worst:.2e}")


def test_k_rows_vs_oracle_screen_and_beta_k(screen2):
    """Flat screen (open surface: boundary vertices in one or two triangles),
    lambda = 0.5; then beta_K of a random vertex datum and the combined
    right-hand side against the oracle."""
    from elastodynamicebem_amd import api
    from oracle import Oracle
    pb = _pb(1.0, 16, lam=0.5)
    o = Oracle(screen2.coordinates, screen2.triangles, rho=1.0, mu=1.0, lam=0.5, Tfin=1.0, nT=16)
    o.set_ghc(16, 3, 64)
    with api.Solver(screen2, pb, 7) as s:
        s.set_inner_rule(16, 3, 64)
        LK = s.k_num_blocks()
        s.assemble_k()
        _check_k_rows(s, o, (0, 4), LK)
        _check_k_rows(s, o, (60, 64), LK)
        rng = np.random.default_rng(11)
        g = rng.normal(size=(3 * screen2.number_nodes, 16))
        b = s.beta_k(g)
        rows = (0, 4)
        bo = o.beta_k(g, rows)[3 * rows[0]:3 * rows[1]]
        assert frob_rel(b[3 * rows[0]:3 * rows[1]], bo) < 1e-12
        # the full beta_K against the GPU's own K blocks (all rows)
        K = [s.get_k_block(l) for l in range(LK)]
        ref = np.zeros_like(b)
        for n in range(16):
            for j in range(n + 1):
                if n - j < LK:
                    ref[:, n] += K[n - j] @ g[:, j]
        assert frob_rel(b, ref) < 1e-13


def _bar_u(h, cP, Tfin, z, t):
    """1-D bar solution of calcBoundDataDirichlet.m:6-13 (numpy restatement)."""
    khat = int(np.ceil(cP * Tfin / (2 * h))) - 1
    k = np.arange(khat + 1)[:, None]
    a = cP * t - 2 * h * k - (h - z)
    b = cP * t - 2 * h * (k + 1) + (h - z)
    return np.sum((-1.0) ** k * (a * (a > 0) - b * (b > 0)), axis=0)


def _bar_dudz(h, cP, Tfin, z, t):
    khat = int(np.ceil(cP * Tfin / (2 * h))) - 1
    k = np.arange(khat + 1)[:, None]
    a = cP * t - 2 * h * k - (h - z)
    b = cP * t - 2 * h * (k + 1) + (h - z)
    return np.sum((-1.0) ** k * ((a > 0) * 1.0 + (b > 0) * 1.0), axis=0)


def test_dirichlet_data_and_rhs():
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd.mesh import read_mesh
    bar = read_mesh(mesh_path("barH1_1"))
    pb = _pb(6.0, 48, rho=0.25, mu=0.5, lam=0.0, domain="barH1")
    cP = np.sqrt((pb.lambda_ + 2 * pb.mu) / pb.rho)
    with api.Solver(bar, pb, 7) as s:
        g = s.dirichlet_datum(api.DATUM_REF_BAR_H1)
        t = 6.0 / 48 * np.arange(1, 49)
        z = bar.coordinates[:, 2]
        ref = np.stack([_bar_u(1.0, cP, 6.0, z, tn) for tn in t], axis=1)
        assert np.abs(g[2::3] - ref).max() < 1e-13 * max(1.0, np.abs(ref).max())
        assert np.all(g[0::3] == 0) and np.all(g[1::3] == 0)
        s.set_inner_rule(16, 3, 64)
        s.assemble_k()
        beta = s.rhs_dirichlet(api.DATUM_REF_BAR_H1)
        r = s.rhs(api.DATUM_REF_BAR_H1)
        bk = s.beta_k(g)
        assert frob_rel(beta, r / 2 + bk) < 1e-15
    # sphereWave datum: the vertex trace has the plus sign, the rhs the minus sign
    sw = read_mesh(mesh_path("sphereWave_1"))
    pb = _pb(12.0, 192, rho=1.0, mu=1.0, lam=2.0, domain="sphereWave")
    with api.Solver(sw, pb, 7) as s:
        g = s.dirichlet_datum(api.DATUM_REF_SPHERE_WAVE)
        tt = 12.0 / 192 * np.arange(1, 193)
        x0 = sw.coordinates[:, 0:1]
        ref = np.exp(-20 * (x0 - 2 + 2.0 * tt[None, :] - 0.475) ** 2)
        assert np.abs(g[0::3] - ref).max() < 1e-13
        r = s.rhs(api.DATUM_REF_SPHERE_WAVE)
        assert r[0::3].min() <= 0 and np.abs(r[0::3]).max() > 0


def _dd_bar(level, nT):
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd.mesh import read_mesh
    bar = read_mesh(mesh_path(f"barH1_{level}"))
    pb = _pb(6.0, nT, rho=0.25, mu=0.5, lam=0.0, domain="barH1")
    pb.lev = level
    with api.Solver(bar, pb, 7) as s:
        s.set_inner_rule(16, 3, 64)
        s.assemble(api.PREC_GHC)
        s.assemble_k()
        s.rhs_dirichlet(api.DATUM_REF_BAR_H1)
        a = s.mot_solve(None)
        tm = s.timings()
    cP = 2.0
    dt = 6.0 / nT
    tmid = dt * (np.arange(nT) + 0.5)
    c = bar.center
    nz = bar.normal[:, 2]
    ex = np.zeros_like(a)
    for n in range(nT):
        ex[2::3, n] = (pb.lambda_ + 2 * pb.mu) * _bar_dudz(1.0, cP, 6.0, c[:, 2], tmid[n]) * nz
    cos = float(np.sum(a * ex) / (np.linalg.norm(a) * np.linalg.norm(ex)))
    rel = frob_rel(a, ex)
    print(f"DD barH1_{level} ({bar.numberTriangles} triangles, {nT} steps): K {tm['k_assembly_ms']:.1f} ms, "
          f"beta_K {tm['beta_k_ms']:.2f} ms, MOT {tm['mot_ms']:.1f} ms; density vs closed-form traction: "
          f"cos {cos:+.4f}, rel L2 {rel:.3e}")
    assert np.all(np.isfinite(a))
    return cos, rel


def test_dd_bar_known_answer():
    """Direct Dirichlet driver on the reference's barH1 problems (rho 0.25,
    mu 0.5, lambda 0, t in [0, 6]), method FN 7 16 3 64: the density is the
    traction of the closed-form solution, sigma n = (lambda + 2 mu) du/dz n_z e_z
    on the end faces and 0 on the lateral faces (taken at mid-step), to the
    discretisation error, which falls under refinement (level 1 with 48 steps
    -> level 2 with 96).  A wrong sign or scale of K, or of the 1/2 of the
    right-hand side, leaves no such agreement."""
    cos1, rel1 = _dd_bar(1, 48)
    assert cos1 > 0.99 and rel1 < 0.15
    cos2, rel2 = _dd_bar(2, 96)
    assert cos2 > 0.99 and rel2 < 0.8 * rel1


def test_k_rows_sphere_wave():
    """The reference's sphereWave problem (inputFiles/input_waveOnSphere.txt:
    rho 1, mu 1, lambda 2 -- the kernelK.cu typo has weight --, t in [0, 12],
    192 steps, 1488 triangles, 746 vertices), rule FN 7 16 3 64: sampled test
    rows of every K_lag and of the direct Dirichlet right-hand side against
    the oracle."""
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd.mesh import read_mesh
    from oracle import Oracle
    m = read_mesh(mesh_path("sphereWave_1"))
    pb = _pb(12.0, 192, rho=1.0, mu=1.0, lam=2.0, domain="sphereWave")
    o = Oracle(m.coordinates, m.triangles, rho=1.0, mu=1.0, lam=2.0, Tfin=12.0, nT=192, nGH=7)
    o.set_ghc(16, 3, 64)
    with api.Solver(m, pb, 7) as s:
        s.set_inner_rule(16, 3, 64)
        LK = s.k_num_blocks()
        assert LK == o.k_num_blocks()
        s.assemble_k()
        info = s.k_store_info()
        worst = _check_k_rows(s, o, (700, 702), LK)
        g = s.dirichlet_datum(api.DATUM_REF_SPHERE_WAVE)
        bk = s.beta_k(g)
        ms = s.timings()["k_assembly_ms"]
    rows = (700, 702)
    bo = o.beta_k(g, rows)[3 * rows[0]:3 * rows[1]]
    assert frob_rel(bk[3 * rows[0]:3 * rows[1]], bo) < 1e-12
    print(f"K sphereWave_1: LK {LK}, {info['slots']} blocks in {ms:.0f} ms, worst lag error {worst:.2e}")


def test_k_per_pair_path_agrees(sphere2, monkeypatch):
    """The per-(test triangle, vertex) form of the regular part
    (EBEM_K_BY_FIELD=0: k_kblocks, the reference's accumulation order) and
    the default by field triangle and level (k_ksf + k_kgather) give the same
    K_lag rows to round-off."""
    from elastodynamicebem_amd import api
    pb = _pb(4.0, 32)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("EBEM_K_BY_FIELD", mode)
        with api.Solver(sphere2, pb, 7) as s:
            s.set_inner_rule(16, 3, 64)
            LK = s.k_num_blocks()
            s.assemble_k()
            out[mode] = [s.get_k_block(l, (100, 104)) for l in range(LK)]
    peak = max(np.linalg.norm(K) for K in out["0"])
    worst = max(np.linalg.norm(a - b) for a, b in zip(out["0"], out["1"])) / peak
    print(f"K per pair vs by field triangle and level, sphere2 rows 100-103: {worst:.2e}")
    assert worst < 1e-12
