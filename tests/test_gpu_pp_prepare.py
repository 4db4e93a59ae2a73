"""ebem_pp_prepare: the density-independent part of a post-processing request
run on the library's second stream while the marching runs, then finished by
ebem_pp_eval.  The field must be the same bit for bit as without preparing,
and every request the preparation does not match must still be evaluated in
full."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(mesh, Tfin, nT):
    from elastodynamicebem_amd import api
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=Tfin, nT=nT)
    return api.Solver(mesh, pb, 7)


def _points(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None] * rng.uniform(1.1, 2.0, size=(n, 1))


def test_prepared_post_processing_is_bit_identical(sphere2):
    from elastodynamicebem_amd import api
    rng = np.random.default_rng(21)
    Tfin, nT = 4.0, 32
    s = _solver(sphere2, Tfin, nT)
    x = _points(rng, 53)
    t = Tfin / nT * np.arange(1, nT + 1)
    s.assemble(api.PREC_FP64)
    s.rhs(api.DATUM_SPHERE_T2EXP)
    s.mot_solve(None)
    u0 = s.pp_eval(x, t)
    assert s.timings()["pp_prepare_ms"] == 0
    # prepared beside the marching
    s.pp_prepare(x, t)
    s.mot_solve(None)
    u1 = s.pp_eval(x, t)
    assert s.timings()["pp_prepare_ms"] > 0
    assert np.array_equal(u0, u1)
    assert np.abs(u0).max() > 0
    # a different request than the prepared one: evaluated in full
    x2 = _points(rng, 17)
    s.pp_prepare(x, t)
    s.mot_solve(None)
    u2 = s.pp_eval(x2, t)
    assert s.timings()["pp_prepare_ms"] == 0
    s.mot_solve(None)
    assert np.array_equal(u2, s.pp_eval(x2, t))
    # a call that may change what the preparation depends on discards it
    s.pp_prepare(x, t)
    s.assemble(api.PREC_FP64)
    s.mot_solve(None)
    u3 = s.pp_eval(x, t)
    assert s.timings()["pp_prepare_ms"] == 0
    assert np.array_equal(u0, u3)
    s.close()


def test_two_time_groups_are_not_prepared(sphere2):
    """Times with two different offsets from the step grid form two groups of
    the convolution: ebem_pp_prepare leaves them to ebem_pp_eval."""
    from elastodynamicebem_amd import api
    rng = np.random.default_rng(22)
    Tfin, nT = 4.0, 32
    dt = Tfin / nT
    s = _solver(sphere2, Tfin, nT)
    x = _points(rng, 9)
    t = np.concatenate([dt * np.arange(1, nT + 1), dt * np.arange(1, 9) - 0.5 * dt])
    s.assemble(api.PREC_FP64)
    s.rhs(api.DATUM_SPHERE_T2EXP)
    s.mot_solve(None)
    u0 = s.pp_eval(x, t)
    s.pp_prepare(x, t)
    s.mot_solve(None)
    u1 = s.pp_eval(x, t)
    assert s.timings()["pp_prepare_ms"] == 0
    assert np.array_equal(u0, u1)
    s.close()
