"""CPU: planted mistakes in the oracle (built with -DORACLE_MUTATION=n) must
fail the same closed-form checks the real oracle passes."""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import closed_forms as cf
from conftest import system
from oracle_py import ROOT, Oracle


def _mutant(n):
    p = os.path.join(ROOT, "oracle", "_mut", f"liboracle_m{n}.so")
    if not os.path.exists(p):
        pytest.fail(f"{p} missing: run make -C oracle")
    return Oracle(p)


def _thomas_ok(o):
    m, jac = system(4, 9, 3, seed=3, shift=6.0)
    Dia, DinvC = o.lineLU(m, jac["jacD"], jac["jacO1"], jac["jacO2"])
    Aout = o.assembleTriBlocks(m, jac["jacO1"], jac["jacO2"])
    R = np.random.default_rng(1).uniform(-1, 1, 3 * m.nelem)
    dU = o.solveDU(m, Dia, DinvC, Aout, R)
    ref = -spla.splu(cf.assemble_M(m, jac).tocsc()).solve(R)
    return np.allclose(dU, ref, rtol=1e-12, atol=1e-14)


def _res_ok(o):
    m, jac = system(3, 5, 4, seed=2)
    U = np.random.default_rng(2).uniform(-1, 1, 4 * m.nelem)
    R = o.lineRes(m, jac["jacD"], jac["jacO1"], jac["jacO2"], U, jac["rhs"])
    return np.allclose(R, jac["rhs"] + cf.assemble_J(m, jac) @ U, rtol=1e-13)


def _gemv_ok(o):
    nv = 5
    rng = np.random.default_rng(4)
    LU = rng.uniform(-1, 1, nv * nv * 3)
    x = rng.uniform(-1, 1, nv * 3)
    b = o.dgemvLU(nv, LU, x)
    for e in range(3):
        F = LU[nv * nv * e: nv * nv * (e + 1)].reshape(nv, nv, order="F")
        ref = (np.tril(F, -1) + np.eye(nv)) @ (np.triu(F) @ x[nv * e: nv * (e + 1)])
        if not np.allclose(b[nv * e: nv * (e + 1)], ref, rtol=1e-13, atol=1e-14):
            return False
    return True


def test_real_oracle_passes(oracle):
    assert _thomas_ok(oracle) and _res_ok(oracle) and _gemv_ok(oracle)


@pytest.mark.parametrize("n,check", [(1, _thomas_ok), (2, _thomas_ok), (3, _thomas_ok), (4, _res_ok), (5, _gemv_ok)])
def test_mutant_fails(n, check):
    o = _mutant(n)
    assert o.lib.orc_mutation() == n
    assert not check(o)
