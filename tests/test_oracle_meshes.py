"""CPU: the mesh variants used by the GPU tests are valid LIJ systems: the
oracle's Thomas solve on them equals the sparse direct solve of M, and its
residual equals the assembled J product."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import closed_forms as cf
import synth_np
from conftest import system
from gpu_util import cross_lines, ragged_lines, reversed_lines

VARIANTS = [
    ("reversed", lambda: reversed_lines(system(5, 13, 9)[0])),
    ("cross", lambda: cross_lines(7, 9, 5)),
    ("ragged", lambda: ragged_lines([1, 2, 17, 64, 3, 100, 1, 65], 9)),
]


@pytest.mark.parametrize("name,make", VARIANTS, ids=[v[0] for v in VARIANTS])
def test_variant_is_consistent(oracle, name, make):
    m = make()
    jac = synth_np.generate(11, m.nvar, m.nelem, m.nintface, 2.0 * m.nvar + 2)
    Dia, DinvC = oracle.lineLU(m, jac["jacD"], jac["jacO1"], jac["jacO2"])
    A = oracle.assembleTriBlocks(m, jac["jacO1"], jac["jacO2"])
    R = np.random.default_rng(5).uniform(-1, 1, m.nvar * m.nelem)
    dU = oracle.solveDU(m, Dia, DinvC, A, R)
    ref = -spla.splu(cf.assemble_M(m, jac).tocsc()).solve(R)
    assert np.allclose(dU, ref, rtol=1e-12, atol=1e-14)
    U = np.random.default_rng(6).uniform(-1, 1, m.nvar * m.nelem)
    Rr = oracle.lineRes(m, jac["jacD"], jac["jacO1"], jac["jacO2"], U, jac["rhs"])
    assert np.allclose(Rr, jac["rhs"] + cf.assemble_J(m, jac) @ U, rtol=1e-13)
    out = oracle.lij_run(m, jac, 6)
    h, _, _ = cf.lij_closed_form(m, jac, 6)
    assert cf.history_close(out["hist"], h)[0]
