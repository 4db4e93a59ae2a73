"""Handle state across a rejected mesh and the surrogate-solve guard."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_rejected_mesh_leaves_no_stale_state(screen2):
    """A valid mesh is assembled, then a degenerate mesh is rejected: the
    handle must not keep the old store or mesh (the next solve fails cleanly
    instead of indexing the old device buffers with the new sizes)."""
    import ctypes as C
    from elastodynamicebem_amd import api, _lib
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=1.0, nT=16)
    with api.Solver(screen2, pb, 7) as s:
        s.assemble(api.PREC_FP64)
        s.rhs(api.DATUM_SCREEN_SIN2)
        s.mot_solve(None)
        X = np.ascontiguousarray(screen2.coordinates[:3].copy())
        X[2] = X[0] + 2 * (X[1] - X[0])  # collinear: zero area
        T = np.array([[0, 1, 2]], np.int32)
        rc = s.lib.ebem_set_mesh(s.h, 1, 3, _lib.dptr(X), T.ctypes.data_as(_lib._ip), 0)
        assert rc == -1  # EBEM_EINVAL
        out = np.zeros(3 * 128 * 16)
        assert s.lib.ebem_mot_solve(s.h, None, _lib.dptr(out)) == -1
        assert s.lib.ebem_assemble(s.h, 0) == -1
        assert b"mesh" in s.lib.ebem_last_error()
        # a bad vertex index is rejected the same way
        T2 = np.array([[0, 1, 7]], np.int32)
        X2 = np.ascontiguousarray(screen2.coordinates[:3])
        assert s.lib.ebem_set_mesh(s.h, 1, 3, _lib.dptr(X2), T2.ctypes.data_as(_lib._ip), 0) == -1
        # and a valid mesh brings the handle back
        coords = np.ascontiguousarray(screen2.coordinates)
        tri = np.ascontiguousarray(screen2.triangles, np.int32)
        _lib.check(s.lib.ebem_set_mesh(s.h, 128, coords.shape[0], _lib.dptr(coords), tri.ctypes.data_as(_lib._ip), 0))
        s.assemble(api.PREC_FP64)
        s.rhs(api.DATUM_SCREEN_SIN2)
        a = s.mot_solve(None)
        assert np.isfinite(a).all() and np.abs(a).max() > 0


def test_solve_on_surrogate_store_needs_opt_in(screen2):
    from elastodynamicebem_amd import api, _lib
    from elastodynamicebem_amd import surrogate as SG
    p = np.zeros(SG.PARAMS, np.float32)
    nm = np.zeros(SG.NORM, np.float32)
    nm[24:48] = 1.0
    nm[66:84] = 1.0
    nm[84:108] = -1e30
    nm[108:132] = 1e30
    sur = SG.Surrogate()
    try:
        sur.set_params(p, nm)
        pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=1.0, nT=16)
        with api.Solver(screen2, pb, 7) as s:
            sur.attach(s)
            s.assemble(api.PREC_SURROGATE)
            s.rhs(api.DATUM_SCREEN_SIN2)
            with pytest.raises(_lib.EbemError, match="surrogate"):
                s.mot_solve(None)
            s.set_option(api.OPT_SURROGATE_SOLVE, 1)
            a = s.mot_solve(None)
            assert np.isfinite(a).all()
            sur.detach(s)
    finally:
        sur.close()
