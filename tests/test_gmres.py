"""Line-preconditioned flexible GMRES: the CPU restatement (oracle/gmres.py)
pinned by closed forms, and the GPU solver (csrc/gmres.hip) against it."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import closed_forms as cf
from conftest import ROOT, system

sys.path.insert(0, os.path.join(ROOT, "oracle"))
from gmres import LinePreconditionedGmres  # noqa: E402


def direct(m, jac):
    J = cf.assemble_J(m, jac).tocsc()
    return J, spla.spsolve(J, -jac["rhs"])


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ---------------------------------------------------------------- oracle pins
@pytest.mark.parametrize("shift,restart", [(12.0, 20), (3.0, 5), (2.0, 30)])
def test_oracle_gmres_reaches_direct_solution(oracle, shift, restart):
    m, jac = system(8, 20, 9, seed=3, shift=shift)
    gm = LinePreconditionedGmres(oracle, m, jac)
    x, hist, it = gm.solve(jac["rhs"], restart=restart, max_iters=80, rtol=1e-10)
    J, xs = direct(m, jac)
    assert hist[-1] <= 1e-10 * hist[0]
    assert rel(x, xs) < 1e-9
    # the Givens estimate is the true residual of the returned U (to round-off)
    true = np.linalg.norm(jac["rhs"] + J @ x)
    assert abs(true - hist[-1]) <= 1e-12 * hist[0]


def test_oracle_gmres_where_lij_diverges(oracle):
    m, jac = system(8, 20, 9, seed=3, shift=2.0)
    assert oracle.lij_run(m, jac, 12)["hist"][-1] > 10 * jac["rhs"].std()   # LIJ alone grows
    x, hist, it = LinePreconditionedGmres(oracle, m, jac).solve(jac["rhs"], restart=30, max_iters=80, rtol=1e-10)
    assert hist[-1] <= 1e-10 * hist[0] and rel(x, direct(m, jac)[1]) < 1e-9


def test_oracle_gmres_exact_preconditioner(oracle):
    """One line and no off-line blocks: the line solve is J^-1, so one
    iteration reaches round-off."""
    m, jac = system(1, 30, 9, seed=5, shift=12.0)
    x, hist, it = LinePreconditionedGmres(oracle, m, jac).solve(jac["rhs"], restart=5, rtol=1e-13)
    assert it == 1 and hist[-1] <= 1e-13 * hist[0]


def test_oracle_gmres_more_sweeps_fewer_iterations(oracle):
    m, jac = system(8, 20, 9, seed=3, shift=3.0)
    gm = LinePreconditionedGmres(oracle, m, jac)
    _, _, it1 = gm.solve(jac["rhs"], restart=30, rtol=1e-10, sweeps=1)
    _, _, it2 = gm.solve(jac["rhs"], restart=30, rtol=1e-10, sweeps=2)
    assert it2 < it1


# ------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def gctx(tbs):
    c = tbs.Context(0, 9)
    yield c
    c.close()


def gpu_gmres(tbs, ctx, m, jac, prec=0, restart=20, max_iters=80, rtol=1e-10, sweeps=1):
    from gpu_util import dev, host, zeros
    p = tbs.Plan(ctx, m, prec)
    p.factor(dev(jac["jacD"]), dev(jac["jacO1"]), dev(jac["jacO2"]))
    g = tbs.Gmres(p, restart)
    U = zeros(m.nvar * m.nelem)
    hist, it = g.solve(dev(jac["rhs"]), U, max_iters, rtol, sweeps)
    st = g.stats()
    g.close()
    p.close()
    return host(U), hist, it, st


# history tolerance: the preconditioners differ by round-off (Gauss-Jordan vs
# LU factors) and so do the dot-product summation orders; a slowly converging
# Krylov sequence (shift 2, where LIJ alone diverges) amplifies that to ~1e-6
@pytest.mark.gpu
@pytest.mark.parametrize("shift,restart,sweeps,htol", [(12.0, 20, 1, 1e-8), (3.0, 5, 1, 1e-7), (2.0, 30, 1, 1e-4),
                                                      (3.0, 30, 2, 1e-7)])
def test_gpu_gmres_matches_oracle(tbs, gctx, oracle, shift, restart, sweeps, htol):
    print(f"case gmres shift {shift} restart {restart} sweeps {sweeps}", flush=True)
    m, jac = system(8, 20, 9, seed=3, shift=shift)
    U, hist, it, st = gpu_gmres(tbs, gctx, m, jac, restart=restart, sweeps=sweeps)
    x, href, itref = LinePreconditionedGmres(oracle, m, jac).solve(jac["rhs"], restart=restart, max_iters=80,
                                                                   rtol=1e-10, sweeps=sweeps)
    assert it == itref
    keep = href > 1e-9 * href[0]
    assert np.allclose(hist[keep], href[keep], rtol=htol, atol=0)
    assert hist[-1] <= 1e-10 * hist[0]
    assert rel(U, direct(m, jac)[1]) < 1e-9 and rel(U, x) < 1e-9
    assert st["true_residual"] <= 1e-10 * hist[0] * 1.01


@pytest.mark.gpu
def test_gpu_gmres_deterministic_and_fp32_preconditioner(tbs, gctx):
    m, jac = system(16, 64, 9, seed=9, shift=2.5)
    a = gpu_gmres(tbs, gctx, m, jac, restart=25)
    b = gpu_gmres(tbs, gctx, m, jac, restart=25)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    c = gpu_gmres(tbs, gctx, m, jac, prec=1, restart=25)        # FP32 factors in the preconditioner
    assert c[1][-1] <= 1e-10 * c[1][0]
    assert rel(c[0], direct(m, jac)[1]) < 1e-9
