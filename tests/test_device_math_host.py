"""The GPU formulation of the SA inner integration (elastodynamicebem_amd/csrc/ebem_sa.hpp,
compiled for the host by tests/native/Makefile) against the oracle, on the CPU.

This checks the re-organised evaluation (folded kernel weights, sparse B-basis,
single inverse rotation) before any GPU run; the GPU tests then check the
device build of the same header through the C ABI."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libsa_host.so")
_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def host_sa():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native")], check=True)
    lib = C.CDLL(LIB)
    lib.sa_point_block.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
    lib.sa_point_block.restype = C.c_int
    return lib


def _run(host_sa, mesh, ntrial, seed, spread, coplanar_every=0):
    from oracle import Oracle
    o = Oracle(mesh.coordinates, mesh.triangles)
    rng = np.random.default_rng(seed)
    errs = []
    for trial in range(ntrial):
        j = int(rng.integers(o.N))
        T = o.coords[o.tri[j]]
        rec = o.field_record(T, o.normal[j])
        c = o.center[int(rng.integers(o.N))] + rng.normal(size=3) * spread * (trial % 3)
        if coplanar_every and trial % coplanar_every == 0:
            c[2] = 0.0
        t = rng.uniform(0.01, 1.0)
        ref = o.pp_block(c, j, t)
        out = np.zeros(9)
        host_sa.sa_point_block(rec.ctypes.data_as(_dp), np.ascontiguousarray(c).ctypes.data_as(_dp), t, o.velP,
                               o.velS, 1.0, out.ctypes.data_as(_dp))
        out = out.reshape(3, 3).T
        out[np.abs(out) < 1e-14] = 0
        sc = np.abs(ref).max()
        if sc == 0 and np.abs(out).max() == 0:
            continue
        errs.append(np.abs(out - ref).max() / max(sc, 1e-300))
    return np.array(errs)


def test_sphere_points(host_sa, sphere2):
    e = _run(host_sa, sphere2, 1500, 0, 0.02)
    assert len(e) > 300
    assert np.median(e) < 1e-13
    assert e.max() < 1e-8


def test_screen_points(host_sa, screen2):
    e = _run(host_sa, screen2, 1500, 1, 0.02, coplanar_every=2)
    assert len(e) > 300
    assert np.median(e) < 1e-13
    # near-coplanar small-zeta points: both evaluations of the reference's
    # closed forms carry ~1e-6 cancellation error (tests/test_oracle.py)
    assert np.quantile(e, 0.99) < 1e-9
    assert e.max() < 1e-5


@pytest.mark.parametrize("fma", [False, True])
def test_host_replay_cfg0_lags(screen2, fma):
    """Replay of the GPU assembly (W_m items, second differences) on the flat
    8x8 screen with the host build of the device math, with and without FMA
    instructions: E_3 and E_9 match the oracle to 1e-12 (Frobenius).  Guards
    against contraction inside the cancellation-prone closed forms."""
    import hostasm
    from oracle import Oracle
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "native")], check=True)
    o = Oracle(screen2.coordinates, screen2.triangles, Tfin=1.0, nT=16)
    E = hostasm.assemble(o, hostasm.load(fma), 16, rows=range(0, 128, 3), lags=[3, 9])
    for l in (3, 9):
        R, _ = o.matrix_block(l)
        G = E[l][0::1]
        rows = np.concatenate([np.arange(3 * i, 3 * i + 3) for i in range(0, 128, 3)])
        assert np.linalg.norm(G[rows] - R[rows]) <= 1e-12 * np.linalg.norm(R[rows]), l
