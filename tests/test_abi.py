"""The C ABI library loads and exports every symbol that include/ebem.h declares."""
import os
import re

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "ebem.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ebem_[A-Za-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    from elastodynamicebem_amd import _lib
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES)


def test_no_gpu_calls_error_cleanly():
    """Without a device, ebem_create fails with an error code and a message (no crash)."""
    import ctypes as C
    import torch
    from elastodynamicebem_amd import _lib
    if torch.cuda.device_count() > 0:
        return
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.ebem_create(C.byref(h), 0)
    assert rc < 0
    assert len(lib.ebem_last_error()) > 0


def test_api_rejects_bad_method():
    import pytest
    from elastodynamicebem_amd import api
    with pytest.raises(ValueError):
        api.BEMenerg_setupCore("SA 05")
    with pytest.raises(NotImplementedError):
        api.BEMenerg_setupCore("MX.G2D 7 64")
    with pytest.raises(ValueError):
        api.BEMenerg_setupCore("FN 7 32 3 256")
    mi, n, w, *_ = api.BEMenerg_setupCore("SA 07")
    assert mi.numNodiExt == 7 and n.shape == (7, 3) and abs(w.sum() - 1) < 1e-12
    mi, n, w, GHCn, GHCw, DIAGn, DIAGw = api.BEMenerg_setupCore("FN 7 64 3 256")
    assert mi.numNodiInt == 192 and GHCn.shape == (192, 3) and DIAGn.shape == (256, 3)
    assert abs(GHCw.sum() - 1) < 1e-12 and abs(DIAGw.sum() - 1) < 1e-14


def test_inner_rules_match_the_oracle():
    """The library's composite and collapsed rules (host code of the shipped
    library) equal the oracle's restatement of GaussHammerComposite.m and
    doppioGauss1D."""
    import numpy as np
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd.mesh import screen_uniform
    from oracle import Oracle
    m = screen_uniform(1)
    o = Oracle(m.coordinates, m.triangles)
    for nsub, nsing, ndiag in ((64, 3, 256), (16, 7, 64), (4, 19, 1024), (1, 1, 16)):
        a, aw, b, bw = api.inner_rule_tables(nsub, nsing, ndiag)
        oa, oaw = o.gh_composite(nsub, nsing)
        ob, obw = o.doppio_gauss(ndiag)
        assert np.abs(a - oa).max() < 1e-15 and np.abs(aw - oaw).max() < 1e-17
        assert np.abs(b - ob).max() < 1e-15 and np.abs(bw - obw).max() < 1e-17


def test_read_input_file(tmp_path):
    from elastodynamicebem_amd import api
    p = tmp_path / "in.txt"
    p.write_text("%rho\n1\n%mu\n1\n%nu\n0.25\n%lambda\n1\n%T\n4\n%nT\n128\n%mesh\nsphereUniform\n%lev\n3\n"
                 "%BIE\nEFIE\n%BOU\nDIR\n")
    pb = api.readInputFile(str(p))
    assert pb.nT == 128 and pb.domainType == "sphereUniform" and abs(pb.velP - 3 ** 0.5) < 1e-15


def test_load_points_tables():
    import numpy as np
    from elastodynamicebem_amd import api
    pb = api.PbParam(domainType="sphereWave", Tfin=4.0)
    x, t, i, n, tp = api.BEMenerg_postProc_loadPoints(pb)
    c = np.linspace(-5, 5, 250)
    assert x.shape[0] == sum(1 for a in c for b in c if a * a + b * b > 1) and n == x.shape[0] * 8
    # column-major order of meshgrid: the first column (x = -5) runs through every y first
    assert np.all(x[:250, 0] == -5) and np.allclose(x[:250, 1], c) and tp == "u(:, t)"
    x, t, i, n, tp = api.BEMenerg_postProc_loadPoints(api.PbParam(domainType="barH1", Tfin=2.0))
    assert x.shape == (3, 3) and t.size == 200 and t[-1] == 2.0 and i == [3]
    x, t, i, n, tp = api.BEMenerg_postProc_loadPoints(api.PbParam(domainType="screenUniform"))
    assert n == 0 and i == [3]


def test_stats_and_timings_need_a_handle():
    """The statistics and timing entry points reject a null handle or output
    with an error code (no device needed)."""
    import ctypes as C
    import numpy as np
    from elastodynamicebem_amd import _lib
    lib = _lib.load()
    out = np.zeros(20)
    st = np.zeros(5, np.int64)
    assert lib.ebem_timings(None, out.ctypes.data_as(_lib._dp), 20) < 0
    assert lib.ebem_assembly_kp_stats(None, st.ctypes.data_as(_lib._i64p)) < 0
    assert lib.ebem_pp_kp_stats(None, st.ctypes.data_as(_lib._i64p)) < 0
    src = open(os.path.join(ROOT, "include", "ebem.h")).read()
    assert "#define EBEM_NUM_TIMINGS 20" in src
    from elastodynamicebem_amd import api
    assert len(api.Solver._KP_KEYS) == 5
    # the double-layer entry points reject a null handle too
    n = C.c_int()
    i64 = C.c_int64()
    assert lib.ebem_k_num_blocks(None, C.byref(n)) < 0
    assert lib.ebem_assemble_k(None) < 0
    assert lib.ebem_k_store_info(None, C.byref(i64), C.byref(i64), C.byref(i64)) < 0
    assert lib.ebem_get_k_block_rows(None, 0, 0, 1, out.ctypes.data_as(_lib._dp)) < 0
    assert lib.ebem_dirichlet_datum(None, 4, out.ctypes.data_as(_lib._dp)) < 0
    assert lib.ebem_beta_k(None, out.ctypes.data_as(_lib._dp), out.ctypes.data_as(_lib._dp)) < 0
    assert lib.ebem_rhs_dirichlet(None, 4, out.ctypes.data_as(_lib._dp)) < 0
    assert lib.ebem_pp_eval_direct(None, None, None, 0, None, 0, None, out.ctypes.data_as(_lib._dp)) < 0
    assert lib.ebem_pp_prepare(None, 0, None, 0, None) < 0
    assert lib.ebem_neumann_datum(None, 4, out.ctypes.data_as(_lib._dp)) < 0
    assert lib.ebem_beta_v(None, out.ctypes.data_as(_lib._dp), out.ctypes.data_as(_lib._dp)) < 0
    assert lib.ebem_mot_solve_neumann(None, 4, out.ctypes.data_as(_lib._dp)) < 0
