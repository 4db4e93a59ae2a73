"""MATLAB .mat outputs (elastodynamicebem_amd/matio.py): the reference's file and
variable names, and round trips of density, lag blocks and fields."""
import numpy as np
import scipy.io as sio
import scipy.sparse as sp


def test_output_names_follow_the_reference():
    from elastodynamicebem_amd import api, matio
    pb = api.PbParam(domainType="sphereUniform", lev=3)
    mi = api.BEMenerg_setupCore("FN 7 64 3 256")[0]
    assert matio.output_base(pb, "FN", mi, "outputData") == "outputData/sphereUniform3_FN_7_64_3_256"
    assert matio.output_base(pb, "NEW", mi, "outputData") == "outputData/sphereUniform3_NEW_7_64_3_256"
    assert matio.output_base(pb, "SA", mi, "outputData") == "outputData/sphereUniform3_SA_7"
    assert matio.output_base(pb, "A", None, "outputData") == "outputData/sphereUniform3_A_campoVett"
    assert matio.output_base(pb, "NEW_PP", None, "outputData") == "outputData/sphereUniform3_NEW_campoVett"


def test_round_trips(tmp_path):
    from elastodynamicebem_amd import matio
    rng = np.random.default_rng(0)
    dens = rng.normal(size=(30, 8))
    p = matio.save_density(str(tmp_path / "x"), dens)
    assert p.endswith("x_density.mat")
    assert np.array_equal(matio.load_density(p), dens)
    raw = sio.loadmat(p)
    assert raw["density"].shape == (30, 8) and raw["density"].dtype == np.float64
    E = [sp.random(30, 30, density=0.1, random_state=k, format="csc") for k in range(4)]
    p = matio.save_matrices(str(tmp_path / "y"), E)
    back = matio.load_matrices(p)
    assert len(back) == 4 and all((a != b).nnz == 0 for a, b in zip(E, back))
    assert sio.loadmat(p)["matrixSaved"].shape == (4, 1)
    p = matio.save_lag_matrix(str(tmp_path / "z_matrixE3"), E[3].toarray())
    assert (matio.load_lag_matrix(p) != E[3]).nnz == 0
    f = rng.normal(size=(5, 7, 3))
    p = matio.save_field(str(tmp_path / "campo"), f)
    assert np.array_equal(matio.load_field(p), f)
    cell = sio.loadmat(p)["campoVett"]
    assert cell.shape == (5, 7) and cell[2, 3].shape == (3, 1)
