import os
import sys

import numpy as np
import pytest

# torch first: its bundled HIP runtime must be the one the process loads, as
# in bench.py; loading the library's system runtime first leaves torch
# without devices.
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

MESHES = os.path.join(ROOT, "tests", "golden", "meshes")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def mesh_path(name):
    return os.path.join(MESHES, name + ".mesh")


@pytest.fixture(scope="session")
def screen2():
    from elastodynamicebem_amd.mesh import read_mesh
    return read_mesh(mesh_path("screenUniform_2"))


@pytest.fixture(scope="session")
def sphere2():
    from elastodynamicebem_amd.mesh import read_mesh
    return read_mesh(mesh_path("sphereUniform_2"))


@pytest.fixture(scope="session")
def sphere3():
    from elastodynamicebem_amd.mesh import read_mesh
    return read_mesh(mesh_path("sphereUniform_3"))


def frob_rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a - b)


def ulp_oracle(mesh, Tfin, nT, nGH=7, seed=0):
    """Oracle on coordinates moved by one ulp (random direction per entry): its
    distance to the plain oracle is the reference algorithm's own sensitivity."""
    from oracle import Oracle
    rng = np.random.default_rng(seed)
    X = mesh.coordinates
    X2 = np.where(rng.random(X.shape) < 0.5, np.nextafter(X, np.inf), np.nextafter(X, -np.inf))
    return Oracle(X2, mesh.triangles, Tfin=Tfin, nT=nT, nGH=nGH)
