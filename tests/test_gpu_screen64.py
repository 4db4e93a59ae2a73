"""Full-size flat screen: [-0.5, 0.5]^2 in 64 x 64 squares (8192 triangles,
24576 DOF), t in [0, 2] with 64 steps, datum (0, 0, sin^2(pi t / 2)), plain
FP64 and mixed precision, on the device.

Sampled test rows of E_lag against the oracle (FP64 bar 1e-10, or 3x the
oracle's own change under a one-ulp perturbation of the coordinates where the
reference algorithm is less reproducible than that: late lags of the corner
rows move by 1e-4 .. 1 under such a perturbation), the store size
against the reference's light-cone count, and size-independent properties of
the solved density: the mesh and the datum are invariant under the rotation by
pi about the screen's normal, so the density is too (x, y components change
sign, z does not).  Mixed precision (FP32 far field) and mixed precision with
the learned near-field surrogate (the network of the 1280-triangle sphere's
test: 2M seeded samples, 20 epochs) are solved on the same problem and their
density errors against the FP64 run printed.  Mixed precision: 8.8e-4
measured, held below 2e-3.  Mixed + surrogate: the marching on this open
screen diverges (measured density error 2.4e25 with 1.3M predicted blocks,
training loss 4.7e-4) although the moment projection keeps it bounded on the
closed sphere (4.8e-2, test_gpu_surrogate_cfg3.py); the test holds that
measured behaviour, so a change to it shows up here."""
import time

import numpy as np
import pytest

from conftest import frob_rel, ulp_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def screen64():
    from elastodynamicebem_amd.mesh import screen_uniform
    return screen_uniform(64)


def _rot_pairs(mesh):
    """Index of the triangle each triangle maps to under (x, y) -> (-x, -y)."""
    c = mesh.center
    key = {tuple(np.round(p[:2], 9)): k for k, p in enumerate(c)}
    return np.array([key[tuple(np.round(-p[:2], 9))] for p in c])


def test_screen64_rows_store_and_density(screen64):
    from elastodynamicebem_amd import api
    from oracle import Oracle
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=2.0, nT=64)
    with api.Solver(screen64, pb, 7) as s:
        assert s.num_blocks() == 47
        t0 = time.perf_counter()
        s.assemble(api.PREC_FP64)
        t_asm = time.perf_counter() - t0
        info = s.store_info()
        assert abs(info["pair_lags"] - 7.565e8) / 7.565e8 < 5e-3, info
        o = Oracle(screen64.coordinates, screen64.triangles, Tfin=2.0, nT=64)
        o2 = ulp_oracle(screen64, 2.0, 64)
        for rows in ((0, 2), (4131, 4133)):
            for lag in (0, 1, 7, 30, 46):
                G = s.get_block_rows(lag, *rows)
                R = o.matrix_block(lag, rows=rows)[0][3 * rows[0]:3 * rows[1]]
                R2 = o2.matrix_block(lag, rows=rows)[0][3 * rows[0]:3 * rows[1]]
                if np.linalg.norm(R) > 0:
                    e, noise = frob_rel(G, R), frob_rel(R2, R)
                else:
                    e, noise = np.linalg.norm(G), np.linalg.norm(R2)
                assert e < max(1e-10, 3 * noise), (rows, lag, e, noise)
        s.rhs(api.DATUM_SCREEN_SIN2)
        a = s.mot_solve(None)
        tm = s.timings()
        import bench
        hist_bytes = bench.history_bytes_blocked(s.lag_blocks(), 64)
        assert np.all(np.isfinite(a)) and np.linalg.norm(a) > 0
        P = _rot_pairs(screen64)
        A = a.reshape(-1, 3, a.shape[1])
        sc = np.abs(A).max()
        assert np.abs(A[:, 2] - A[P, 2]).max() < 1e-9 * sc
        assert np.abs(A[:, :2] + A[P, :2]).max() < 1e-9 * sc
        s.assemble(api.PREC_MIXED)
        st = s.assembly_stats()
        am = s.mot_solve(None)
        tmm = s.timings()
        # mixed precision + surrogate (the combined form of the 64 x 64 screen)
        from elastodynamicebem_amd import surrogate as SG
        sur = SG.Surrogate()
        try:
            ntr, _ = SG.build_training_set(sur, 2_000_000, seed=0, held_frac=0.02)
            sur.init(ntr, seed=0)
            losses = sur.train(ntr, 20, batch=2048, lr=2e-3, seed=0)
            sur.attach(s)
            s.set_option(api.OPT_SURROGATE_SOLVE, 1)
            s.assemble(api.PREC_MIXED | api.PREC_SURROGATE)
            tms = s.timings()
            ss = s.surrogate_stats()
            ams = s.mot_solve(None)
            sur.detach(s)
        finally:
            sur.close()
    err = frob_rel(am, a)
    err_ms = frob_rel(ams, a) if np.isfinite(ams).all() else float("inf")
    print(f"screen64: assembly {t_asm:.2f} s wall ({tm['levels_ms']:.0f} ms integrals), "
          f"{info['pair_lags'] / (tm['assembly_ms'] * 1e-3):.3e} pair.lags/s; "
          f"MOT {tm['mot_ms']:.0f} ms (history {tm['history_ms']:.0f} ms, "
          f"{hist_bytes / (tm['history_ms'] * 1e-3) / 1e9:.0f} GB/s); mixed assembly "
          f"{tmm['assembly_ms']:.0f} ms, {st['far_items']} FP32 items, density rel L2 vs FP64 {err:.3e}; "
          f"mixed + surrogate assembly {tms['assembly_ms']:.0f} ms, {ss['blocks']} predicted blocks, "
          f"training loss {losses[0]:.3e} -> {losses[-1]:.3e}, density rel L2 vs FP64 {err_ms:.3e}")
    assert err < 2e-3
    assert ss["blocks"] > 1_000_000
    assert err_ms > 1.0  # unstable on this open screen (see the module docstring)
