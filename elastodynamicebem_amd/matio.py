"""MATLAB .mat outputs of the reference's drivers, written and read with
scipy.io (MAT v5, MATLAB's default `save` format for variables below 2 GB).

File names and variable names follow the reference:
  * SA driver: outputData/<domainType><lev>_SA_<numNodiExt>_density.mat,
    variable `density` (3N x nT) (SA/BEMenerg_coreSA_timeMarching.m:32,116-117);
    the lag blocks tempData/<...>_SA_<numNodiExt>_matrixE<lag>.mat, variable
    `matrix` (sparse 3N x 3N) (:31,42-43,97-98);
  * numeric (FN) driver: outputData/<...>_FN_<ext>_<sub>_<sing>_<diag>_density.mat
    and _matrix.mat with the cell array `matrixSaved` of sparse lag blocks
    (FN/BEMenerg_coreFN_timeMarching.m:33-35,174-180);
  * direct driver: outputData/<...>_NEW_<ext>_<sub>_<sing>_<diag>_density.mat
    (dir/BEMenerg_dir_timeMarchingD.m:34-36,300-301);
  * post-processing: outputData/<domainType><lev>_A_campoVett.mat
    (ppA/BEMenerg_postProcA_setup.m:21-22) and _NEW_campoVett.mat
    (dirPP/BEMenerg_ppD_setup.m:23-24), variable `campoVett`, an
    xSize x tSize cell array of 3 x 1 vectors.
MATLAB's `save` appends .mat to a name without extension; so do these
writers.  The readers use scipy.io.loadmat, which only parses data."""
from __future__ import annotations

import os

import numpy as np
import scipy.io as sio
import scipy.sparse as sp


def _num(v) -> str:
    """num2str of an integer-valued number."""
    return str(int(v))


def output_base(pbParam, form: str, methodInfo=None, root: str = "outputData") -> str:
    """Path stem the reference's driver of `form` writes its results under:
    form "SA" (analytic), "FN" (numeric inner quadrature), "NEW" (direct
    driver), "A" or "NEW_PP" (post-processing fields)."""
    stem = f"{pbParam.domainType}{_num(pbParam.lev)}"
    if form == "SA":
        tail = f"_SA_{_num(methodInfo.numNodiExt)}"
    elif form in ("FN", "NEW"):
        tail = (f"_{form}_{_num(methodInfo.numNodiExt)}_{_num(methodInfo.numSubRegion)}_"
                f"{_num(methodInfo.numNodiSing)}_{_num(methodInfo.numNodiDiag)}")
    elif form == "A":
        tail = "_A_campoVett"
    elif form == "NEW_PP":
        tail = "_NEW_campoVett"
    else:
        raise ValueError(f"unknown output form {form!r}")
    return os.path.join(root, stem + tail)


def _path(p: str) -> str:
    return p if p.endswith(".mat") else p + ".mat"


def save_density(base: str, density: np.ndarray) -> str:
    """<base>_density.mat with variable `density` (3N x nT, double)."""
    p = _path(base + "_density")
    os.makedirs(os.path.dirname(p) or ".", exist_ok=True)
    sio.savemat(p, {"density": np.asarray(density, dtype=np.float64)}, do_compression=False)
    return p


def load_density(path: str) -> np.ndarray:
    return np.asarray(sio.loadmat(_path(path))["density"], dtype=np.float64)


def save_lag_matrix(path: str, E) -> str:
    """One lag block as the variable `matrix` (sparse 3N x 3N)."""
    p = _path(path)
    os.makedirs(os.path.dirname(p) or ".", exist_ok=True)
    sio.savemat(p, {"matrix": sp.csc_matrix(E)}, do_compression=False)
    return p


def load_lag_matrix(path: str) -> sp.csc_matrix:
    return sp.csc_matrix(sio.loadmat(_path(path))["matrix"])


def save_matrices(base: str, blocks) -> str:
    """<base>_matrix.mat with the cell array `matrixSaved` (one sparse matrix per lag)."""
    p = _path(base + "_matrix")
    os.makedirs(os.path.dirname(p) or ".", exist_ok=True)
    cell = np.empty((len(blocks), 1), dtype=object)
    for k, E in enumerate(blocks):
        cell[k, 0] = sp.csc_matrix(E)
    sio.savemat(p, {"matrixSaved": cell}, do_compression=False)
    return p


def load_matrices(path: str) -> list:
    cell = sio.loadmat(_path(path))["matrixSaved"]
    return [sp.csc_matrix(cell[k, 0]) for k in range(cell.shape[0])]


def save_field(base: str, campoVett: np.ndarray) -> str:
    """<base>.mat with the cell array `campoVett` (xSize x tSize of 3 x 1);
    campoVett as returned by the post-processing functions: (xSize, tSize, 3)."""
    f = np.asarray(campoVett, dtype=np.float64)
    p = _path(base)
    os.makedirs(os.path.dirname(p) or ".", exist_ok=True)
    cell = np.empty(f.shape[:2], dtype=object)
    for i in range(f.shape[0]):
        for k in range(f.shape[1]):
            cell[i, k] = f[i, k].reshape(3, 1)
    sio.savemat(p, {"campoVett": cell}, do_compression=False)
    return p


def load_field(path: str) -> np.ndarray:
    cell = sio.loadmat(_path(path))["campoVett"]
    out = np.zeros(cell.shape + (3,))
    for i in range(cell.shape[0]):
        for k in range(cell.shape[1]):
            out[i, k] = np.asarray(cell[i, k], dtype=np.float64).ravel()
    return out
