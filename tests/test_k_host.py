"""The device formulation of the double-layer operator K
(elastodynamicebem_amd/csrc/ebem_k_math.hpp, compiled for the host by
tests/native/Makefile: the same k_reg_lane / k_sing_lane as the kernel, the 64
lanes summed by the same butterfly) against the oracle, on the CPU, before
any GPU run.  FP64 bar 1e-10 on the scale of the pair's largest block (the
singular part cancels large terms; late blocks are round-off around zero)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libk_host.so")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


@pytest.fixture(scope="module")
def host_k():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native")], check=True)
    lib = C.CDLL(LIB)
    lib.k_block_host.argtypes = ([C.c_int, C.c_int, _dp, _ip, _dp, _dp, _dp] + [C.c_double] * 4 +
                                 [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int,
                                  _dp])
    lib.k_block_host.restype = C.c_int
    lib.ppp_block_host.argtypes = ([C.c_int, C.c_int, _dp, _ip, _dp, _dp] + [C.c_double] * 4 +
                                   [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int, _dp])
    lib.ppp_block_host.restype = C.c_int
    lib.k_chunk_vs_single_host.argtypes = ([C.c_int, C.c_int, _dp, _ip, _dp, _dp, _dp] + [C.c_double] * 4 +
                                           [C.c_int, C.c_int, _dp, _dp, C.c_int, _dp] + [C.c_int] * 4 + [_dp])
    lib.k_chunk_vs_single_host.restype = C.c_int
    lib.k_host_set_sf_mode.argtypes = [C.c_int]
    return lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _host_block(lib, o, rules, lam, lag, S, s):
    rin, rdg = rules
    out = np.zeros(9)
    lib.k_block_host(o.N, o.coords.shape[0], _p(o.coords), o.tri.ctypes.data_as(_ip), _p(o.normal), _p(o.area),
                     _p(o.maxL), 1.0, 1.0, lam, o.Tfin, o.nT, o.nGH, _p(o.ghn), _p(o.ghw), rin.shape[0], _p(rin),
                     rdg.shape[0], _p(rdg), lag, S, s, _p(out))
    return out.reshape(3, 3).T.copy()


@pytest.mark.parametrize("mode", [0, 1], ids=["per_vertex", "by_field_levels"])
@pytest.mark.parametrize("meshname,lam", [("sphere2", 2.0), ("screen2", 0.5)])
def test_host_k_blocks_match_oracle(host_k, request, meshname, lam, mode):
    """mode 0: the regular part per (S, s) pair lag by lag (k_reg_lane);
    mode 1: by field triangle as the device's default path, level sums of
    every F != S with vertex s combined -1, 3, -3, 1 (k_sf_lane)."""
    from elastodynamicebem_amd import api
    from oracle import Oracle
    mesh = request.getfixturevalue(meshname)
    o = Oracle(mesh.coordinates, mesh.triangles, rho=1.0, mu=1.0, lam=lam, Tfin=4.0, nT=32)
    o.set_ghc(16, 3, 64)
    gn, gw, dn, dw = api.inner_rule_tables(16, 3, 64)
    rules = (np.ascontiguousarray(np.column_stack([gn, gw])), np.ascontiguousarray(np.column_stack([dn, dw])))
    assert np.allclose(gn, o.ghcn) and np.allclose(gw, o.ghcw) and np.allclose(dn, o.dgn) and np.allclose(dw, o.dgw)
    host_k.k_host_set_sf_mode(mode)
    rng = np.random.default_rng(7)
    worst = 0.0
    checked = 0
    for S in rng.choice(o.N, 5, replace=False):
        S = int(S)
        verts = list(o.tri[S])
        nb = sorted({int(v) for F in range(o.N) if len(set(o.tri[F]) & set(verts)) > 0 for v in o.tri[F]})
        cand = verts + [v for v in nb if v not in verts][:2] + [int(rng.integers(o.coords.shape[0]))]
        for s in cand:
            blocks = [(l, o.k_block(l, S, s)) for l in range(0, 14)]
            peak = max(np.linalg.norm(b) for _, b in blocks)
            for l, R in blocks:
                G = _host_block(host_k, o, rules, lam, l, S, s)
                if peak == 0:
                    assert np.abs(G).max() == 0
                    continue
                e = np.linalg.norm(G - R) / peak
                worst = max(worst, e)
                assert e < 1e-10, (S, s, l, e)
                checked += 1
    host_k.k_host_set_sf_mode(0)
    assert checked > 100
    print(f"worst host-vs-oracle K block error {worst:.2e} over {checked} blocks (mode {mode})")


def test_host_pp_p_blocks_match_oracle(host_k, sphere2):
    """The device math of the direct post-processing's double-layer part
    (ebem_ppd_math.hpp) against the oracle's kernelP restatement."""
    from elastodynamicebem_amd import api
    from oracle import Oracle
    o = Oracle(sphere2.coordinates, sphere2.triangles, rho=1.0, mu=1.0, lam=2.0, Tfin=8.0, nT=32)
    o.set_pp_rule(16, 3)
    gn, gw, _, _ = api.inner_rule_tables(16, 3, 16)
    rule = np.ascontiguousarray(np.column_stack([gn, gw]))
    assert np.allclose(gn, o.ppn) and np.allclose(gw, o.ppw)
    rng = np.random.default_rng(8)
    checked = 0
    worst = 0.0
    for _ in range(150):
        x = rng.normal(size=3) * rng.choice([0.3, 0.9, 1.6])
        s = int(rng.integers(sphere2.number_nodes))
        t = rng.uniform(0.2, 8.0)
        y = int(rng.integers(int(np.ceil(t / o.dt))))
        R = o.pp_p_block(x, t, y, s)
        out = np.zeros(18)
        rc = host_k.ppp_block_host(o.N, sphere2.number_nodes, _p(o.coords), o.tri.ctypes.data_as(_ip),
                                   _p(o.normal), _p(o.area), 1.0, 1.0, 2.0, 8.0, 32, rule.shape[0], _p(rule),
                                   _p(np.ascontiguousarray(x)), t, y, s, _p(out))
        assert rc == 0, "chunked and per-block double-layer sums differ"
        G = out[:9].reshape(3, 3).T  # per vertex (k_ppd_P)
        Gf = out[9:].reshape(3, 3).T  # by field triangle and level (k_ppf, k_ppf_gather)
        sc = np.abs(R).max()
        if sc == 0:
            assert np.abs(G).max() == 0 and np.abs(Gf).max() == 0
            continue
        assert np.abs(G - R).max() < 1e-11 * sc, (x, t, y, s)
        assert np.abs(Gf - R).max() < 1e-11 * sc, (x, t, y, s)
        worst = max(worst, np.abs(Gf - R).max() / sc)
        checked += 1
    assert checked > 40
    print(f"double-layer PP blocks: field-triangle form vs oracle worst {worst:.2e} over {checked} blocks")


def test_chunked_regular_sum_is_bitwise_the_per_lag_sum(host_k, sphere2):
    """The kernel's chunked regular sum (each level once for up to four
    consecutive lags) gives every lag exactly the per-lag sum."""
    from elastodynamicebem_amd import api
    from oracle import Oracle
    o = Oracle(sphere2.coordinates, sphere2.triangles, rho=1.0, mu=1.0, lam=2.0, Tfin=4.0, nT=32)
    gn, gw, _, _ = api.inner_rule_tables(16, 3, 16)
    rin = np.ascontiguousarray(np.column_stack([gn, gw]))
    rng = np.random.default_rng(9)
    for _ in range(12):
        S = int(rng.integers(o.N))
        s = int(o.tri[int(rng.integers(o.N))][int(rng.integers(3))])
        lagA = int(rng.integers(0, 20))
        nA = int(rng.integers(1, 5))
        out = np.zeros(18 * nA)
        assert host_k.k_chunk_vs_single_host(o.N, sphere2.number_nodes, _p(o.coords), o.tri.ctypes.data_as(_ip),
                                             _p(o.normal), _p(o.area), _p(o.maxL), 1.0, 1.0, 2.0, 4.0, 32, 7,
                                             _p(o.ghn), _p(o.ghw), rin.shape[0], _p(rin), lagA, nA, S, s,
                                             _p(out)) == 0
        assert np.array_equal(out[:9 * nA], out[9 * nA:]), (S, s, lagA, nA)


def test_bench_k_op_counts_match_the_counting_build():
    """The per-point operation counts behind the K roofline in bench.py equal
    the op-counting host build of the device math (every double arithmetic
    operation and sqrt counted once)."""
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    import bench
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native")], check=True)
    lib = C.CDLL(os.path.join(HERE, "native", "libk_opcount.so"))
    lib.k_opcount_sf_point.argtypes = [C.c_int, C.c_int]
    lib.k_opcount_sf_point.restype = C.c_longlong
    lib.k_opcount_sing_level.restype = C.c_longlong
    for mA in range(1, 9):
        for nL in range(1, bench.K_SF_CHUNK + 1):
            assert int(bench.k_sf_point_flops(mA, nL)) == lib.k_opcount_sf_point(mA, nL), (mA, nL)
    # one outer node, one collapsed-rule node in each of the three children
    assert 3 * bench.K_SING_POINT_LEVEL == lib.k_opcount_sing_level()
