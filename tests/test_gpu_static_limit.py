"""Telescoping static limit of the device assembly: sum_l (nT - l) E_l equals
the elastostatic Kelvin single-layer Galerkin matrix once c_S T exceeds every
outer-node / field-vertex distance (tests/test_oracle.py states the identity
and pins the oracle with it; here the HIP path meets the same independent
target).  The 320-triangle sphere, lambda = 2, mu = 1, rho = 1.5, T = 4,
64 steps (c_S T = 3.27 > the diameter 2)."""
import numpy as np
import pytest

from test_oracle import kelvin_block

pytestmark = pytest.mark.gpu


def test_device_static_limit_is_kelvin(sphere2):
    from elastodynamicebem_amd import api
    from oracle import Oracle
    lam, mu, rho, Tfin, nT = 2.0, 1.0, 1.5, 4.0, 64
    pb = api.PbParam(rho=rho, mu=mu, lambda_=lam, Tfin=Tfin, nT=nT)
    rows = (0, 3)
    with api.Solver(sphere2, pb, 7) as s:
        s.assemble(api.PREC_FP64)
        L = s.num_blocks()
        S = sum((nT - l) * s.get_block_rows(l, *rows) for l in range(L))
    o = Oracle(sphere2.coordinates, sphere2.triangles, Tfin=Tfin, nT=nT, rho=rho, mu=mu, lam=lam)
    worst = 0.0
    for i in range(*rows):
        for j in (0, 1, 2, 9, 40, 161, 250, 319):
            K = kelvin_block(o, sphere2, i, j, lam, mu)
            got = S[3 * (i - rows[0]):3 * (i - rows[0]) + 3, 3 * j:3 * j + 3]
            worst = max(worst, np.abs(got - K).max() / np.abs(K).max())
    print(f"device static limit vs Kelvin: worst {worst:.2e}")
    assert worst < 1e-9
