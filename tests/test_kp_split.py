"""kP ranges of the assembly (k_levels_kp), on the CPU through the host build
of the device math (tests/native/sa_host.cpp).

Between the fronts, while the S sphere of every outer node is short of the
field plane and the P sphere has passed the whole field triangle, each child
of the fan has the single edge region of coeffG_TR_P.m whose coefficients do
not depend on t, so W(t) = WP + t^2 WQ and E_l = 2 dt^2 WQ / (4 pi rho) for the
lags whose three levels lie in the range.  Checked here:
  * the oracle's E_l on those lags equals that closed form (the oracle's W is
    quadratic in t there, a property of the reference algorithm that a dropped
    or mis-signed t^2 term would break);
  * the split W_m equals the per-level W_m of k_levels to rounding;
  * E built from the split W agrees with the oracle as closely as E built from
    the per-level W does.
Sphere of 320 triangles, T = 4, 128 steps (c_P = sqrt 3, c_S = 1)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT

_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def setup():
    from elastodynamicebem_amd.mesh import read_mesh
    from hostasm import exact_windows, load
    from oracle import Oracle
    lib = load()
    lib.sa_item_kp.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp,
                               C.c_int]
    lib.sa_item_kp.restype = C.c_int
    lib.sa_kp_range.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.sa_item_bf.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                               C.c_int]
    lib.sa_item_bf.restype = C.c_int
    m = read_mesh(os.path.join(ROOT, "tests", "golden", "meshes", "sphereUniform_2.mesh"))
    o = Oracle(m.coordinates, m.triangles, Tfin=4.0, nT=128)
    lo, hi = exact_windows(o, o.num_blocks())
    return lib, o, lo, hi


def _gh(o, i):
    V = o.coords[o.tri[i]]
    nodes = (o.ghn[:, 0:1] * V[0][None, :] + o.ghn[:, 1:2] * V[1][None, :]) + o.ghn[:, 2:3] * V[2][None, :]
    gh = np.zeros(128)
    n = o.nGH
    gh[0:n], gh[32:32 + n], gh[64:64 + n] = nodes[:, 0], nodes[:, 1], nodes[:, 2]
    gh[96:96 + n] = o.area[i] * o.ghw
    return gh


def _blk(e, inv):
    e = e * inv
    e[np.abs(e) < 1e-14] = 0
    b = np.zeros((3, 3))
    for q, (r, c) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
        b[r, c] = b[c, r] = e[q]
    return b


def test_kp_split_matches_levels_and_oracle(setup):
    lib, o, lo, hi = setup
    N, dt = o.N, o.dt
    inv = 1.0 / (4 * np.pi * o.rho)
    rng = np.random.default_rng(3)
    found = tried = 0
    worst_w = worst_closed = worst_split = 0.0
    while found < 25 and tried < 5000:
        tried += 1
        i, j = (int(x) for x in rng.integers(0, N, 2))
        if lo[i, j] > hi[i, j]:
            continue
        gh = _gh(o, i)
        rec = o.field_record(o.coords[o.tri[j]], o.normal[j])
        vF = np.ascontiguousarray(o.coords[o.tri[j]].ravel())
        ka, kb, sa, sb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib.sa_kp_range(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), vF.ctypes.data_as(_dp), o.nGH, o.velP,
                        o.velS, dt, int(lo[i, j]), int(hi[i, j]), C.byref(ka), C.byref(kb), C.byref(sa), C.byref(sb))
        a, b = ka.value, kb.value
        if a > b:
            continue
        found += 1
        assert b - a + 1 >= 3 and max(1, lo[i, j] - 1) <= a and b <= hi[i, j] + 1
        WP, WQ = np.zeros(6), np.zeros(6)
        ok = lib.sa_item_kp(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), o.nGH, a * dt, b * dt, o.velP, o.velS,
                            WP.ctypes.data_as(_dp), WQ.ctypes.data_as(_dp), 0)
        assert ok == 1, (i, j, a, b)
        Wl, Ws = {}, {}
        for m in range(max(1, lo[i, j] - 1), hi[i, j] + 2):
            w = np.zeros(6)
            lib.sa_item(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), o.nGH, m * dt, o.velP, o.velS,
                        w.ctypes.data_as(_dp), 0)
            Wl[m] = w
            Ws[m] = WP + ((m * dt) * (m * dt)) * WQ if a <= m <= b else w
            if a <= m <= b:
                worst_w = max(worst_w, np.abs(Ws[m] - w).max() / np.abs(w).max())
        closed = _blk(2 * dt * dt * WQ, inv)
        for lag in range(max(lo[i, j], a - 1), min(hi[i, j], b + 1) + 1):
            R = o.pair_block(lag, i, j)[0]
            sc = np.abs(R).max()
            if sc == 0:
                continue

            def E(W):
                e = 1.0 * W[lag + 1]
                if lag >= 1:
                    e = e + (-2.0) * W[lag]
                if lag >= 2:
                    e = e + W[lag - 1]
                return _blk(e, inv)
            if a <= lag - 1 and lag + 1 <= b:
                worst_closed = max(worst_closed, np.abs(closed - R).max() / sc)
            e_split, e_lev = np.abs(E(Ws) - R).max() / sc, np.abs(E(Wl) - R).max() / sc
            worst_split = max(worst_split, e_split)
            assert e_split <= max(1e-9, 2 * e_lev), (i, j, lag, e_split, e_lev)
    print(f"kP ranges: {found} pairs; split W vs per-level {worst_w:.1e}; oracle E vs 2 dt^2 WQ {worst_closed:.1e}; "
          f"split E vs oracle {worst_split:.1e}")
    assert found >= 20
    assert worst_w < 1e-13
    assert worst_closed < 1e-9


def test_kp_range_none_when_coplanar(setup):
    """Outer nodes in the field plane (zeta = 0) are never in class kP."""
    lib, o, lo, hi = setup
    from elastodynamicebem_amd.mesh import read_mesh
    from oracle import Oracle
    from hostasm import exact_windows
    m = read_mesh(os.path.join(ROOT, "tests", "golden", "meshes", "screenUniform_2.mesh"))
    s = Oracle(m.coordinates, m.triangles, Tfin=2.0, nT=64)
    l2, h2 = exact_windows(s, s.num_blocks())
    for i in range(s.N):
        gh = _gh(s, i)
        for j in range(s.N):
            if l2[i, j] > h2[i, j]:
                continue
            rec = s.field_record(s.coords[s.tri[j]], s.normal[j])
            vF = np.ascontiguousarray(s.coords[s.tri[j]].ravel())
            ka, kb, sa, sb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            lib.sa_kp_range(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), vF.ctypes.data_as(_dp), s.nGH, s.velP,
                            s.velS, s.dt, int(l2[i, j]), int(h2[i, j]), C.byref(ka), C.byref(kb), C.byref(sa),
                            C.byref(sb))
            assert ka.value > kb.value


def test_pp_point_kp_form_matches_oracle(setup):
    """Post-processing kernel M(x, tau) of ppA/BEMenerg_postProcA_calcMatrixBlock.m
    for one point between the fronts in class kP: WP + tau^2 WQ from
    node_eval_kp (one node, weight 1) against the oracle's M at each tau."""
    lib, o, lo, hi = setup
    rng = np.random.default_rng(5)
    d = rng.normal(size=(40, 3))
    X = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(1.2, 2.0, size=(40, 1))
    inv = 1.0 / (4 * np.pi * o.rho)
    n_ranges = 0
    worst = 0.0
    for x in X:
        for j in rng.integers(0, o.N, 8):
            j = int(j)
            V = o.coords[o.tri[j]]
            dmax = np.sqrt(((V - x) ** 2).sum(1)).max()
            zeta = abs(np.dot(x - V[0], o.normal[j]))
            ta, tb = dmax * (1 + 1e-6) / o.velP, zeta * (1 - 1e-6) / o.velS
            if tb <= ta:
                continue
            gh = np.zeros(128)
            gh[0], gh[32], gh[64], gh[96] = x[0], x[1], x[2], 1.0
            rec = o.field_record(V, o.normal[j])
            WP, WQ = np.zeros(6), np.zeros(6)
            ok = lib.sa_item_kp(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), 1, ta, tb, o.velP, o.velS,
                                WP.ctypes.data_as(_dp), WQ.ctypes.data_as(_dp), 0)
            assert ok == 1
            n_ranges += 1
            for tau in np.linspace(ta, tb, 4):
                R = o.pp_block(x, j, tau)
                G = _blk(WP + tau * tau * WQ, inv)
                worst = max(worst, np.abs(G - R).max() / np.abs(R).max())
    print(f"{n_ranges} point-triangle kP ranges, worst rel {worst:.1e}")
    assert n_ranges >= 50
    assert worst < 1e-10


@pytest.mark.parametrize("fp32", [0, 1])
@pytest.mark.parametrize("which", ["sphere", "screen"])
def test_between_front_ranges_bit_identical(setup, which, fp32):
    """Between-front ranges (k_levels_bf): node_eval_bf in chunks of levels gives
    bit for bit the W_m of the level-by-level evaluation (node_eval), on the
    sphere (classes kP and kSPP, zeta > 0) and on a flat screen (coplanar
    branch of coeffG_TC.m); in FP64, and with the region integrals in FP32 (the
    far-field ranges of the mixed-precision form) against node_eval<float>."""
    lib, o, lo, hi = setup
    if which == "screen":
        from elastodynamicebem_amd.mesh import screen_uniform
        from hostasm import exact_windows
        from oracle import Oracle
        m = screen_uniform(8)
        o = Oracle(m.coordinates, m.triangles, Tfin=2.0, nT=64)
        lo, hi = exact_windows(o, o.num_blocks())
    rng = np.random.default_rng(17)
    ranges = levels = 0
    for _ in range(4000):
        i, j = (int(x) for x in rng.integers(0, o.N, 2))
        if lo[i, j] > hi[i, j]:
            continue
        gh = _gh(o, i)
        rec = o.field_record(o.coords[o.tri[j]], o.normal[j])
        vF = np.ascontiguousarray(o.coords[o.tri[j]].ravel())
        ka, kb, sa, sb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib.sa_kp_range(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), vF.ctypes.data_as(_dp), o.nGH, o.velP,
                        o.velS, o.dt, int(lo[i, j]), int(hi[i, j]), C.byref(ka), C.byref(kb), C.byref(sa),
                        C.byref(sb))
        a, b = sa.value, sb.value
        if a > b:
            continue
        assert b - a + 1 >= 2 and max(1, lo[i, j] - 1) <= a and b <= hi[i, j] + 1
        assert ka.value > kb.value or a == kb.value + 1
        n = b - a + 1
        W = np.zeros(6 * n)
        assert lib.sa_item_bf(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), o.nGH, a, n, o.dt, o.velP, o.velS,
                              W.ctypes.data_as(_dp), fp32) == 1
        for k in range(n):
            w = np.zeros(6)
            lib.sa_item(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), o.nGH, (a + k) * o.dt, o.velP, o.velS,
                        w.ctypes.data_as(_dp), fp32)
            assert np.array_equal(w.view(np.int64), W[6 * k:6 * k + 6].view(np.int64)), (i, j, a + k)
        ranges += 1
        levels += n
    print(f"{which}, {'FP32' if fp32 else 'FP64'} regions: {ranges} between-front ranges, {levels} levels "
          f"bit-identical")
    assert ranges >= 100


def test_kp_split_fp32_regions(setup):
    """kP ranges with the region integrals in FP32 (far-field pairs of the
    mixed-precision form): the same state checks as FP64 (they stay FP64), and
    W = WP + t^2 WQ within FP32 rounding of the FP64 W -- no worse than the
    per-level FP32 evaluation (node_eval<float>) of the same levels, since WP
    and WQ are the coefficients of that evaluation grouped by level weight."""
    lib, o, lo, hi = setup
    N, dt = o.N, o.dt
    rng = np.random.default_rng(5)
    found = tried = 0
    worst_split = worst_lev = 0.0
    while found < 40 and tried < 8000:
        tried += 1
        i, j = (int(x) for x in rng.integers(0, N, 2))
        if lo[i, j] > hi[i, j]:
            continue
        gh = _gh(o, i)
        rec = o.field_record(o.coords[o.tri[j]], o.normal[j])
        vF = np.ascontiguousarray(o.coords[o.tri[j]].ravel())
        ka, kb, sa, sb = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib.sa_kp_range(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), vF.ctypes.data_as(_dp), o.nGH, o.velP,
                        o.velS, dt, int(lo[i, j]), int(hi[i, j]), C.byref(ka), C.byref(kb), C.byref(sa), C.byref(sb))
        a, b = ka.value, kb.value
        if a > b:
            continue
        found += 1
        coef = {}
        for fp32 in (0, 1):
            WP, WQ = np.zeros(6), np.zeros(6)
            ok = lib.sa_item_kp(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), o.nGH, a * dt, b * dt, o.velP,
                                o.velS, WP.ctypes.data_as(_dp), WQ.ctypes.data_as(_dp), fp32)
            assert ok == 1, (i, j, a, b, fp32)
            coef[fp32] = (WP, WQ)
        for m in range(a, b + 1):
            t2 = (m * dt) * (m * dt)
            w64, w32 = np.zeros(6), np.zeros(6)
            lib.sa_item(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), o.nGH, m * dt, o.velP, o.velS,
                        w64.ctypes.data_as(_dp), 0)
            lib.sa_item(rec.ctypes.data_as(_dp), gh.ctypes.data_as(_dp), o.nGH, m * dt, o.velP, o.velS,
                        w32.ctypes.data_as(_dp), 1)
            s64 = coef[0][0] + t2 * coef[0][1]
            s32 = coef[1][0] + t2 * coef[1][1]
            sc = np.abs(w64).max()
            assert np.abs(s64 - w64).max() <= 1e-13 * sc
            worst_split = max(worst_split, np.abs(s32 - w64).max() / sc)
            worst_lev = max(worst_lev, np.abs(w32 - w64).max() / sc)
    print(f"kP ranges, FP32 regions: {found} pairs; W vs FP64: split {worst_split:.1e}, per level {worst_lev:.1e}")
    assert found >= 30
    # measured 2.8e-3 for both: the FP32 closed forms lose 3-4 digits of W
    # (differences of nearly equal terms), split or per level alike; a wrong
    # or dropped term would differ at O(1)
    assert 0 < worst_split < 1e-2
    assert worst_split <= 4 * worst_lev + 1e-6
