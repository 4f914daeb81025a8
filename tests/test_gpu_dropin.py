"""GPU: the drop-in kernels (one per active reference OKL kernel) against the
oracle on identical inputs, through the C ABI."""
import numpy as np
import pytest

import closed_forms as cf
import synth_np
from conftest import system
from gpu_util import cross_lines, dev, host, mctx, ragged_lines, reversed_lines, zeros

pytestmark = pytest.mark.gpu

RTOL = 1e-12   # FP64 parity bound for single kernels (normwise, relative)


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def meshes():
    out = []
    m, _ = system(60, 64, 9)
    out.append(("cfg1_60x64_nv9", m))
    out.append(("reversed_5x13_nv9", reversed_lines(system(5, 13, 9)[0])))
    out.append(("cross_7x9_nv5", cross_lines(7, 9, 5)))
    out.append(("ragged_nv9", ragged_lines([1, 2, 17, 64, 3, 100, 1, 65], 9)))
    out.append(("grid_4x6_nv3", system(4, 6, 3)[0]))
    out.append(("grid_3x20_nv16", system(3, 20, 16)[0]))
    # block sizes past a 16-lane row: the paper's block-size study runs to ~30
    out.append(("grid_3x9_nv17", system(3, 9, 17)[0]))
    out.append(("ragged_nv22", ragged_lines([1, 7, 12, 3], 22)))
    out.append(("grid_2x6_nv32", system(2, 6, 32)[0]))
    out.append(("grid_2x5_nv1", system(2, 5, 1)[0]))
    return out


MESHES = meshes()


@pytest.fixture(scope="module")
def ctxs(tbs):
    cache = {}

    def get(nv):
        if nv not in cache:
            cache[nv] = tbs.Context(0, nv)
        return cache[nv]
    yield get
    for c in cache.values():
        c.close()


def _jac(m, seed=11, shift=None):
    return synth_np.generate(seed, m.nvar, m.nelem, m.nintface, shift if shift is not None else 2.0 * m.nvar + 2)


@pytest.mark.parametrize("name,m", MESHES, ids=[x[0] for x in MESHES])
def test_kernels_match_oracle(ctxs, oracle, name, m):
    print(f"case {name}", flush=True)
    ctx = ctxs(m.nvar)
    jac = _jac(m)
    d = m.to_device(mctx())
    D, O1, O2 = dev(jac["jacD"]), dev(jac["jacO1"]), dev(jac["jacO2"])
    nv, n = m.nvar, m.nelem
    # copies / add: exact
    B = dev(np.zeros(nv * nv * n))
    ctx.copyAtoBjac(n, D, B)
    assert np.array_equal(host(B), jac["jacD"])
    u = np.random.default_rng(0).uniform(-1, 1, nv * n)
    v = np.random.default_rng(1).uniform(-1, 1, nv * n)
    du, dv = dev(u), dev(v)
    ctx.addAtoB(n, du, dv)
    assert np.array_equal(host(dv), v + u)
    ctx.copyAtoB(n, du, dv)
    assert np.array_equal(host(dv), u)
    # A packing: exact
    Ag = dev(np.zeros(nv * nv * n))
    ctx.assembleTriBlocks(d, O1, O2, Ag)
    A_ref = oracle.assembleTriBlocks(m, jac["jacO1"], jac["jacO2"])
    assert np.array_equal(host(Ag), A_ref)
    # lineLU
    Dia_ref, DinvC_ref = oracle.lineLU(m, jac["jacD"], jac["jacO1"], jac["jacO2"])
    Dg = dev(jac["jacD"].copy())
    Cg = dev(np.zeros(nv * nv * n))
    ctx.lineLU(n, m.nintface, d, Dg, O1, O2, Cg)
    assert relerr(host(Dg), Dia_ref) < RTOL
    assert relerr(host(Cg), DinvC_ref) < RTOL
    # dgemvLU on the oracle's factors: the oracle's b = L U x, and for the
    # first element of every line (D'_0 = LU(D_0)) the raw block product D x
    xg = np.random.default_rng(3).uniform(-1, 1, nv * n)
    bg = dev(np.zeros(nv * n))
    ctx.dgemvLU(n, dev(Dia_ref), dev(xg), bg)
    b_ref = oracle.dgemvLU(nv, Dia_ref, xg)
    assert relerr(host(bg), b_ref) < RTOL
    bh = host(bg)
    for l in range(min(m.nline, 8)):
        e = m.lines[m.linepoint[l]]
        Dblk = jac["jacD"][nv * nv * e: nv * nv * (e + 1)].reshape(nv, nv, order="F")
        want = Dblk @ xg[nv * e: nv * (e + 1)]
        assert relerr(bh[nv * e: nv * (e + 1)], want) < 1e-13
    # solveDU on the oracle's factors
    R = np.random.default_rng(2).uniform(-1, 1, nv * n)
    dU_ref = oracle.solveDU(m, Dia_ref, DinvC_ref, A_ref, R)
    dUg = dev(np.zeros(nv * n))
    ctx.solveDU(n, m.nintface, d, dev(Dia_ref), dev(DinvC_ref), dev(A_ref), dUg, dev(R))
    assert relerr(host(dUg), dU_ref) < RTOL
    # lineRes
    Rg = dev(jac["rhs"].copy())
    ctx.lineRes(n, m.nintface, d, D, O1, O2, du, Rg)
    R_ref = oracle.lineRes(m, jac["jacD"], jac["jacO1"], jac["jacO2"], u, jac["rhs"])
    assert relerr(host(Rg), R_ref) < RTOL
    # closed form for the residual too
    R_cf = jac["rhs"] + cf.assemble_J(m, jac) @ u
    assert relerr(host(Rg), R_cf) < RTOL


def run_dropin_loop(ctx, m, jac, iters):
    """src/main.cxx:180-229 with the drop-in kernels; returns (hist, U)."""
    d = m.to_device(mctx())
    nv, n = m.nvar, m.nelem
    D, O1, O2, rhs = dev(jac["jacD"]), dev(jac["jacO1"]), dev(jac["jacO2"]), dev(jac["rhs"])
    z = zeros
    DLU, DinvC, Am, U, dU, R = z(nv * nv * n), z(nv * nv * n), z(nv * nv * n), z(nv * n), z(nv * n), z(nv * n)
    mctx().sync()
    ctx.assembleTriBlocks(d, O1, O2, Am)
    ctx.copyAtoBjac(n, D, DLU)
    ctx.lineLU(n, m.nintface, d, DLU, O1, O2, DinvC)
    ctx.copyAtoB(n, rhs, R)
    hist = [ctx.norm2(nv * n, R)]
    for _ in range(iters):
        ctx.solveDU(n, m.nintface, d, DLU, DinvC, Am, dU, R)
        ctx.addAtoB(n, dU, U)
        ctx.copyAtoB(n, rhs, R)
        ctx.lineRes(n, m.nintface, d, D, O1, O2, U, R)
        hist.append(ctx.norm2(nv * n, R))
    ctx.sync()
    return np.array(hist), host(U)


@pytest.mark.parametrize("NL,NK,shift", [(60, 64, 20.0), (60, 64, 12.0)])
def test_reference_loop_history(ctxs, oracle, NL, NK, shift):
    """cfg1: 10 sweeps of the reference loop; residual history within 1e-12 relative
    (normalised history, floor 1e-4) and U within 1e-12 of the oracle."""
    m, jac = system(NL, NK, 9, seed=1, shift=shift)
    ref = oracle.lij_run(m, jac, 10)
    hist, U = run_dropin_loop(ctxs(9), m, jac, 10)
    ok, worst = cf.history_close(hist, ref["hist"])
    print(f"history worst {worst:.3e}, U relerr {relerr(U, ref['U']):.3e}")
    assert ok, worst
    assert relerr(U, ref["U"]) < RTOL
