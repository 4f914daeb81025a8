"""GPU: the benchmark shapes.  cfg2 (360 x 384) against the full oracle; cfg4
(1200 x 1280, shift 12) through size-independent properties: convergence to
1e-10 at the predicted rate, FP32-factor vs FP64 agreement, drop-in kernel
path vs fused path agreement, bitwise repeatability."""
import os

import numpy as np
import pytest

import closed_forms as cf
from gpu_util import dev, host, mctx, zeros

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_cfg2_against_oracle(tbs, oracle):
    print("case cfg2 360x384 fp64", flush=True)
    ctx = tbs.Context(0, 9)
    m = tbs.structured_mesh(360, 384)
    jac = tbs.generate_host(2, 9, m.nelem, m.nintface, 20.0)
    nthreads = min(16, os.cpu_count() or 1)
    ref = oracle.lij_run(m, jac, 8, nthreads=nthreads)
    plan = tbs.Plan(ctx, m)
    plan.factor(dev(jac["jacD"]), dev(jac["jacO1"]), dev(jac["jacO2"]))
    U = zeros(9 * m.nelem)
    hist, ns = plan.solve(dev(jac["rhs"]), U, 8, 0.0)
    ok, worst = cf.history_close(hist, ref["hist"])
    print(f"cfg2 history worst {worst:.3e} U relerr {relerr(host(U), ref['U']):.3e}")
    assert ok, worst
    assert relerr(host(U), ref["U"]) < 1e-12
    plan.close()
    ctx.close()


def test_cfg4_properties(tbs):
    print("case cfg4 1200x1280 shift 12", flush=True)
    ctx = tbs.Context(0, 9)
    m = tbs.structured_mesh(1200, 1280)
    n, nif = m.nelem, m.nintface
    z = zeros
    D, O1, O2, rhs = z(81 * n), z(81 * nif), z(81 * nif), z(9 * n)
    ctx.generate_device(4, n, nif, 12.0, D, O1, O2, rhs)
    ctx.sync()
    # device generator == host generator on a slice (counter-based streams)
    hs = tbs.generate_host(4, 9, 1000, 1000, 12.0)
    assert np.array_equal(D.numpy(count=81000), hs["jacD"]) and np.array_equal(rhs.numpy(count=9000), hs["rhs"])
    res = {}
    for prec in (0, 1):
        plan = tbs.Plan(ctx, m, prec)
        plan.factor(D, O1, O2)
        U = zeros(9 * n)
        hist, ns = plan.solve(rhs, U, 50, 1e-10)
        res[prec] = (hist, ns, host(U))
        print(f"prec {prec}: sweeps {ns} final {hist[-1] / hist[0]:.3e}", flush=True)
        plan.close()
    h64, n64, U64 = res[0]
    h32, n32, U32 = res[1]
    assert h64[-1] <= 1e-10 * h64[0] and n64 <= 8
    assert h32[-1] <= 1e-10 * h32[0] and n32 <= n64 + 1
    rates = h64[1:4] / h64[:3]
    assert np.all((rates > 0.01) & (rates < 0.04)), rates
    assert relerr(U32, U64) < 1e-9
    # repeatability (FP64)
    plan = tbs.Plan(ctx, m, 0)
    plan.factor(D, O1, O2)
    U = zeros(9 * n)
    hist, ns = plan.solve(rhs, U, 50, 1e-10)
    assert np.array_equal(hist, h64) and np.array_equal(host(U), U64)
    plan.close()
    del U
    # drop-in kernel loop (the reference's launch sequence) agrees with the fused path
    d = m.to_device(mctx())
    DLU, DinvC, Am = z(81 * n), z(81 * n), z(81 * n)
    Ud, dU, R = zeros(9 * n), z(9 * n), z(9 * n)
    ctx.assembleTriBlocks(d, O1, O2, Am)
    ctx.copyAtoBjac(n, D, DLU)
    ctx.lineLU(n, nif, d, DLU, O1, O2, DinvC)
    ctx.copyAtoB(n, rhs, R)
    hd = [ctx.norm2(9 * n, R)]
    for _ in range(n64):
        ctx.solveDU(n, nif, d, DLU, DinvC, Am, dU, R)
        ctx.addAtoB(n, dU, Ud)
        ctx.copyAtoB(n, rhs, R)
        ctx.lineRes(n, nif, d, D, O1, O2, Ud, R)
        hd.append(ctx.norm2(9 * n, R))
    ok, worst = cf.history_close(np.array(hd), h64)
    print(f"drop-in vs fused history worst {worst:.3e}, U relerr {relerr(host(Ud), U64):.3e}")
    assert ok, worst
    assert relerr(host(Ud), U64) < 1e-12
    ctx.close()


def test_norm_past_one_pass(tbs):
    """2,201,600 elements = 34,400 residual groups, more partials than the norm
    kernel sums in one pass of its threads (32 x 1024): ||R_0|| = ||rhs|| (U = 0)
    and the last history entry = ||R|| of the returned residual, in numpy."""
    print("case norm 1720x1280", flush=True)
    ctx = tbs.Context(0, 9)
    m = tbs.structured_mesh(1720, 1280)
    n, nif = m.nelem, m.nintface
    assert (n + 63) // 64 > 32 * 1024
    D, O1, O2, rhs = zeros(81 * n), zeros(81 * nif), zeros(81 * nif), zeros(9 * n)
    ctx.generate_device(8, n, nif, 12.0, D, O1, O2, rhs)
    plan = tbs.Plan(ctx, m, 0)
    plan.factor(D, O1, O2)
    U, R = zeros(9 * n), zeros(9 * n)
    hist, ns = plan.solve(rhs, U, 2, 0.0, R_out=R)
    plan.close()
    assert ns == 2
    assert abs(hist[0] - np.linalg.norm(host(rhs))) <= 1e-13 * hist[0]
    assert abs(hist[-1] - np.linalg.norm(host(R))) <= 1e-13 * hist[-1]
    ctx.close()
