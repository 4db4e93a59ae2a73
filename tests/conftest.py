import os
import sys

import numpy as np
import pytest

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
