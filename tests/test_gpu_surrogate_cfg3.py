"""The learned near-field surrogate on its own workload: the 1280-triangle
sphere (sphereUniform_3), T = 4, 128 steps, datum n t^2 e^{-4t}, field at the
4096 seeded shell points; the network trained here on 2M seeded near-field
(pair, lag) samples of generated icospheres (levels 2-4) and screens
(8-32 subdivisions), FP64 analytic labels, 20 epochs (about 20 s).

What it shows (each number printed):
* with the moment projection (default) the solve is stable and its error
  against the plain FP64 form is bounded: measured density 4.8e-2, field
  8.3e-3 relative L2;
* the raw predictions (projection off) break sum_l E_l = 0 per pair and the
  marching diverges (measured density error 5e7);
* against the network-free predictor (all weights zero: the training mean,
  through the same projection) the network lowers the density error by
  five orders of magnitude (measured 7.4e3 -> 4.8e-2);
* the assembly time it saves over the analytic routine it replaces.
"""
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_surrogate_cfg3_solve_error_and_ablation():
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd import surrogate as SG
    from elastodynamicebem_amd.mesh import read_mesh
    from test_gpu_mixed_cfg2 import _shell_points
    m = read_mesh(os.path.join(ROOT, "tests", "golden", "meshes", "sphereUniform_3.mesh"))
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=4.0, nT=128)
    x = _shell_points()
    t = pb.Tfin / pb.nT * np.arange(1, pb.nT + 1)
    rel = lambda p, q: float(np.linalg.norm(p - q) / np.linalg.norm(q))
    sur = SG.Surrogate()
    try:
        ntr, nhe = SG.build_training_set(sur, 2_000_000, seed=0, held_frac=0.02)
        sur.init(ntr, seed=0)
        losses = sur.train(ntr, 20, batch=2048, lr=2e-3, seed=0)
        assert losses[-1] < 0.1 * losses[0]
        with api.Solver(m, pb, 7) as s:
            def run(prec):
                s.assemble(prec)
                ms = s.timings()["assembly_ms"]
                s.rhs(api.DATUM_SPHERE_T2EXP)
                a = s.mot_solve(None)
                return ms, a, s.pp_eval(x, t)
            t64, a64, u64 = run(api.PREC_FP64)
            sur.attach(s)
            s.set_option(api.OPT_SURROGATE_SOLVE, 1)
            ts, a_s, u_s = run(api.PREC_SURROGATE)
            st = s.surrogate_stats()
            s.set_option(api.OPT_SURROGATE_PROJECT, 0)
            _, a_raw, _ = run(api.PREC_SURROGATE)
            s.set_option(api.OPT_SURROGATE_PROJECT, 1)
            p, nm = sur.params()
            sur.set_params(np.zeros_like(p), nm)
            _, a_mean, _ = run(api.PREC_SURROGATE)
            sur.detach(s)
    finally:
        sur.close()
    e_s, f_s = rel(a_s, a64), rel(u_s, u64)
    e_raw = rel(a_raw, a64) if np.isfinite(a_raw).all() else float("inf")
    e_mean = rel(a_mean, a64) if np.isfinite(a_mean).all() else float("inf")
    print(f"surrogate: {st['blocks']} predicted blocks in {st['ms']:.1f} ms; assembly {ts:.1f} ms vs "
          f"{t64:.1f} ms analytic; density {e_s:.3e}, field {f_s:.3e}; without projection {e_raw:.3e}; "
          f"network-free predictor {e_mean:.3e}")
    assert st["blocks"] > 1_000_000
    assert e_s < 0.1 and f_s < 0.02
    assert e_raw > 1e3
    assert e_mean > 1e3 * e_s
