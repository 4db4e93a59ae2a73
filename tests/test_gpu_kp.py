"""Level ranges on the device: the assembly with the kP ranges (k_levels_kp)
and the between-front ranges (k_levels_bf) (default) against the same
assembly with every level through k_levels (EBEM_KP_SPLIT=0), on the
1280-triangle sphere, T = 4, 128 steps.  Every lag matrix agrees to 1e-12
(Frobenius-relative, absolute against ||E_0|| for lags below 1e-6 ||E_0||);
the ranges cover a large share of the items and no range falls back to the
level-by-level evaluation.  Also the per-level items by outer node (default)
against one item per lane.  Parity of the default path with the oracle is in
test_gpu_parity.py (the screens there take the coplanar branch)."""
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def _assemble(split, node_tasks=True, reuse=True):
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd.mesh import read_mesh
    m = read_mesh(os.path.join(ROOT, "tests", "golden", "meshes", "sphereUniform_3.mesh"))
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=4.0, nT=128)
    env = {"EBEM_KP_SPLIT": "1" if split else "0", "EBEM_NODE_TASKS": "1" if node_tasks else "0",
           "EBEM_NODE_REUSE": reuse if isinstance(reuse, str) else ("1" if reuse else "0")}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        s = api.Solver(m, pb, 7)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    s.assemble(api.PREC_FP64)
    return s


def _worst_lag_difference(a, b):
    """Largest per-lag difference of two assemblies: Frobenius-relative, and
    absolute against ||E_0|| for lags below 1e-6 ||E_0||."""
    e0, worst = None, (0.0, -1)
    for lag in range(a.num_blocks()):
        G, R = a.get_block(lag), b.get_block(lag)
        nr = np.linalg.norm(R)
        if e0 is None:
            e0 = nr
        e = np.linalg.norm(G - R) / (nr if nr >= 1e-6 * e0 else e0)
        worst = max(worst, (e, lag))
    return worst


def test_node_tasks_match_item_per_lane():
    """The per-level items evaluated by outer node (node tasks sorted by front
    state, default) against one item per lane (EBEM_NODE_TASKS=0), kP and
    between-front ranges off so that every level goes through them: the two
    add an item's node integrals in a different order only, so every lag
    matrix agrees to 1e-12 and the solved densities to 1e-12."""
    from elastodynamicebem_amd import api
    a = _assemble(False, node_tasks=True)
    b = _assemble(False, node_tasks=False)
    try:
        assert a.store_info() == b.store_info()
        ta, tb = a.timings()["levels_ms"], b.timings()["levels_ms"]
        e, lag = _worst_lag_difference(a, b)
        print(f"per-level integrals by node {ta:.1f} ms vs by item {tb:.1f} ms; worst lag difference {e:.2e} (lag {lag})")
        assert e < 1e-12, (lag, e)
        a.rhs(api.DATUM_SPHERE_T2EXP)
        b.rhs(api.DATUM_SPHERE_T2EXP)
        da, db = a.mot_solve(None), b.mot_solve(None)
        ed = np.linalg.norm(da - db) / np.linalg.norm(db)
        print(f"density difference {ed:.2e}")
        assert ed < 1e-12
    finally:
        a.close()
        b.close()


def test_static_node_reuse_is_bit_identical():
    """Static outer nodes (both fronts beyond every child's edge) take the
    value of their first static level at the later levels of a pair
    (default) instead of being evaluated again (EBEM_NODE_REUSE=0): the two
    assemblies of the 1280-triangle sphere are equal bit for bit, a large
    share of the node tasks is reused, and no chunk failed the device-side
    check of the reused values."""
    a = _assemble(True, reuse=True)
    b = _assemble(True, reuse=False)
    try:
        assert a.store_info() == b.store_info()
        ta, tb = a.timings(), b.timings()
        print(f"per-level integrals {ta['levels_ms']:.1f} ms with reuse ({ta['node_tasks']:.0f} tasks evaluated, "
              f"{ta['node_tasks_reused']:.0f} reused) vs {tb['levels_ms']:.1f} ms ({tb['node_tasks']:.0f} tasks)")
        assert ta["node_reuse_fallbacks"] == 0
        assert ta["node_tasks_reused"] > 0.1 * tb["node_tasks"]
        assert ta["node_tasks"] + ta["node_tasks_reused"] == tb["node_tasks"]
        for lag in range(a.num_blocks()):
            G, R = a.get_block(lag), b.get_block(lag)
            assert np.array_equal(G, R), lag
    finally:
        a.close()
        b.close()



def test_static_node_reuse_fallback_path():
    """The fallback after a failed check of the reused static values, forced
    from the host (EBEM_NODE_REUSE=fallback: every chunk's check reports a
    failure): each chunk is keyed again without reuse, so the assembly equals
    the one without reuse bit for bit and no task is reused."""
    a = _assemble(True, reuse="fallback")
    b = _assemble(True, reuse=False)
    try:
        assert a.store_info() == b.store_info()
        ta, tb = a.timings(), b.timings()
        print(f"fallbacks {ta['node_reuse_fallbacks']:.0f}; tasks {ta['node_tasks']:.0f} vs {tb['node_tasks']:.0f}")
        assert ta["node_reuse_fallbacks"] >= 1
        assert ta["node_tasks_reused"] == 0 and ta["node_tasks"] == tb["node_tasks"]
        for lag in range(a.num_blocks()):
            assert np.array_equal(a.get_block(lag), b.get_block(lag)), lag
    finally:
        a.close()
        b.close()

def test_kp_split_matches_per_level_assembly():
    a = _assemble(True)
    b = _assemble(False)
    try:
        st = a.assembly_kp_stats()
        stb = b.assembly_kp_stats()
        info = a.store_info()
        L = a.num_blocks()
        assert L == b.num_blocks() and info == b.store_info()
        tm = a.timings()
        ta = tm["levels_ms"] + tm["levels_kp_ms"] + tm["levels_bf_ms"]
        tb = b.timings()["levels_ms"]
        print(f"kP ranges {st['ranges']}, levels {st['levels']}; between-front ranges {st['bf_ranges']}, levels "
              f"{st['bf_levels']}; of {info['items']} items; fallbacks {st['fallbacks']}; integral kernels "
              f"{ta:.1f} ms (per level {tm['levels_ms']:.1f}, kP {tm['levels_kp_ms']:.1f}, between-front "
              f"{tm['levels_bf_ms']:.1f}) vs {tb:.1f} ms per level")
        assert all(v == 0 for v in stb.values())
        assert st["fallbacks"] == 0
        assert st["levels"] > 0.15 * info["items"] and st["bf_levels"] > 0.15 * info["items"]
        e0 = None
        worst = 0.0
        for lag in range(L):
            G, R = a.get_block(lag), b.get_block(lag)
            nr = np.linalg.norm(R)
            if e0 is None:
                e0 = nr
            e = np.linalg.norm(G - R) / (nr if nr >= 1e-6 * e0 else e0)
            worst = max(worst, e)
            assert e < 1e-12, (lag, e)
        print(f"worst lag difference {worst:.2e}")
    finally:
        a.close()
        b.close()


def test_pp_kp_split_matches_per_m_evaluation():
    """Post-processing of the solved density at 1024 points of the shell
    1.2 <= |x| <= 2 and all 128 steps, kP ranges on against off: the fields agree
    to 1e-12 (relative L2), and the ranges cover a large share of the items."""
    import bench
    from elastodynamicebem_amd import api
    a = _assemble(True)
    b = _assemble(False)
    try:
        a.rhs(api.DATUM_SPHERE_T2EXP)
        dens = a.mot_solve(None)
        x = bench.shell_points(1024)
        t = 4.0 / 128 * np.arange(1, 129)
        ua = a.pp_eval(x, t, dens)
        st = a.pp_kp_stats()
        ta = a.timings()["pp_ms"]
        ub = b.pp_eval(x, t, dens)
        tb = b.timings()["pp_ms"]
        print(f"post-processing kP ranges {st['ranges']}, items {st['levels']}; between-front ranges "
              f"{st['bf_ranges']}, items {st['bf_levels']}; fallbacks {st['fallbacks']}; {ta:.1f} ms vs {tb:.1f} ms")
        assert all(v == 0 for v in b.pp_kp_stats().values())
        assert st["ranges"] > 0 and st["bf_ranges"] > 0 and st["fallbacks"] == 0
        e = np.linalg.norm(ua - ub) / np.linalg.norm(ub)
        print(f"field difference {e:.2e}")
        assert e < 1e-12
    finally:
        a.close()
        b.close()


def test_screen_between_front_ranges_bit_identical():
    """Flat 16 x 16 screen (512 triangles), T = 2, 64 steps: every outer node lies
    in the field plane, so there are no kP ranges and the between-front ranges
    take the coplanar branch; their results are bit-identical to the
    level-by-level evaluation, so every lag matrix must equal the
    EBEM_KP_SPLIT=0 assembly exactly, and so must the solved density."""
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd.mesh import screen_uniform
    m = screen_uniform(16)
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=2.0, nT=64)
    sol = []
    for split in (True, False):
        old = os.environ.get("EBEM_KP_SPLIT")
        os.environ["EBEM_KP_SPLIT"] = "1" if split else "0"
        try:
            s = api.Solver(m, pb, 7)
        finally:
            if old is None:
                del os.environ["EBEM_KP_SPLIT"]
            else:
                os.environ["EBEM_KP_SPLIT"] = old
        s.assemble(api.PREC_FP64)
        sol.append(s)
    a, b = sol
    try:
        st = a.assembly_kp_stats()
        print(f"screen16: between-front ranges {st['bf_ranges']}, levels {st['bf_levels']} of "
              f"{a.store_info()['items']} items; kP ranges {st['ranges']}")
        assert st["ranges"] == 0 and st["bf_ranges"] > 0 and st["fallbacks"] == 0
        for lag in range(a.num_blocks()):
            assert np.array_equal(a.get_block(lag), b.get_block(lag)), lag
        a.rhs(api.DATUM_SCREEN_SIN2)
        b.rhs(api.DATUM_SCREEN_SIN2)
        assert np.array_equal(a.mot_solve(None), b.mot_solve(None))
    finally:
        a.close()
        b.close()


def test_pp_ranges_off_grid_times():
    """Post-processing at times off the time grid (t = (k + 0.37) dt and a few
    random ones: several groups of shared offsets delta), 256 shell points,
    ranges on against off: fields agree to 1e-12 and the ranges are used."""
    import bench
    from elastodynamicebem_amd import api
    a = _assemble(True)
    b = _assemble(False)
    try:
        a.rhs(api.DATUM_SPHERE_T2EXP)
        dens = a.mot_solve(None)
        x = bench.shell_points(256, seed=3)
        dt = 4.0 / 128
        rng = np.random.default_rng(4)
        t = np.concatenate([dt * (np.arange(0, 127) + 0.37), rng.uniform(0.1, 3.9, size=9)])
        ua = a.pp_eval(x, t, dens)
        st = a.pp_kp_stats()
        ub = b.pp_eval(x, t, dens)
        e = np.linalg.norm(ua - ub) / np.linalg.norm(ub)
        print(f"off-grid times: kP ranges {st['ranges']}, between-front ranges {st['bf_ranges']}, "
              f"fallbacks {st['fallbacks']}; field difference {e:.2e}")
        assert st["ranges"] > 0 and st["bf_ranges"] > 0 and st["fallbacks"] == 0
        assert e < 1e-12
    finally:
        a.close()
        b.close()
