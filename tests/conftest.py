import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle_py import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def tbs():
    import triblocksolve_amd as t
    t.load()
    return t


def system(NL, NK, nvar=9, seed=1, shift=20.0):
    """Structured mesh + seeded Jacobian (numpy generator, independent of the product)."""
    import synth_np
    from triblocksolve_amd.solver import HostMesh
    mesh = HostMesh(nvar=nvar, **synth_np.structured(NL, NK))
    jac = synth_np.generate(seed, nvar, mesh.nelem, mesh.nintface, shift)
    return mesh, jac
