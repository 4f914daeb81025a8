"""GPU: the benchmark workloads themselves against the FP64 oracle.

  cfg4  1,200 x 1,280 = 1,536,000 elements, 9x9 blocks, shift 12, seed 4 (the
        metric's system): FP64 plan, 8 sweeps from zero, against the oracle's 8
        sweeps; FP32 factors run to 1e-10 against the oracle's solution at the
        round-off floor.
  cfg3  the 64-system, 384,000-element batch (batch_shapes seed 3): every
        system's FP64 history and U against the oracle, FP32-factor U against
        the converged oracle solution, and a sample solved alone bitwise.
  cfg5  a seeded Poisson stream of 24 requests (log-uniform 3,840..384,000
        elements, line lengths 64..960, one above 200,000) through the
        continuous-batching server with FP32 factors and tolerance 1e-10, every
        result against the oracle; then an FP64 server on part of the stream,
        history and U against the oracle's same sweeps.

Tolerances: FP64 history |h_k/h_0 - h'_k/h'_0| <= 1e-12 max(h'_k/h'_0, 1e-4)
and ||U - U'|| <= 1e-12 ||U'|| (h', U' = oracle, same sweeps); FP32 factors
||U - U*|| <= 1e-9 ||U*|| (U* = oracle after 12 sweeps, at the round-off
floor) and h_k/h_0 <= 1e-10 at exit."""
import os

import numpy as np
import pytest

import closed_forms as cf
from gpu_util import dev, host, zeros

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def ctx(tbs):
    c = tbs.Context(0, 9)
    yield c
    c.close()


def test_cfg4_against_oracle(tbs, ctx, oracle):
    print("case cfg4 1200x1280 shift 12 seed 4: oracle", flush=True)
    m = tbs.structured_mesh(1200, 1280)
    n = m.nelem
    jac = tbs.generate_host(4, 9, n, m.nintface, 12.0)
    ref = oracle.lij_run(m, jac, 8, nthreads=NTHREADS)
    print("case cfg4: upload", flush=True)
    D, O1, O2, rhs = dev(jac["jacD"]), dev(jac["jacO1"]), dev(jac["jacO2"]), dev(jac["rhs"])
    del jac
    print("case cfg4: fp64 plan, 8 sweeps", flush=True)
    plan = tbs.Plan(ctx, m, tbs.TBS_FP64)
    plan.factor(D, O1, O2)
    U = zeros(9 * n)
    hist, ns = plan.solve(rhs, U, 8, 0.0)
    plan.close()
    ok, worst = cf.history_close(hist, ref["hist"])
    e64 = relerr(host(U), ref["U"])
    print(f"cfg4 fp64: history worst {worst:.3e}, U relerr {e64:.3e}", flush=True)
    assert ns == 8 and ok, worst
    assert e64 <= 1e-12
    print("case cfg4: fp32 factors to 1e-10", flush=True)
    plan = tbs.Plan(ctx, m, tbs.TBS_FP32)
    plan.factor(D, O1, O2)
    U32 = zeros(9 * n)
    h32, n32 = plan.solve(rhs, U32, 50, 1e-10)
    plan.close()
    e32 = relerr(host(U32), ref["U"])
    print(f"cfg4 fp32 factors: {n32} sweeps, final {h32[-1] / h32[0]:.3e}, U relerr vs oracle {e32:.3e}", flush=True)
    assert h32[-1] <= 1e-10 * h32[0]
    assert e32 <= 1e-9
    for x in (D, O1, O2, rhs, U, U32):
        x.close()


def test_cfg3_batch_against_oracle(tbs, ctx, oracle):
    print("case cfg3 batch of 64 systems, 384000 elements", flush=True)
    shapes = tbs.batch_shapes(3, nsys=64, total=384000, lmin=96, lmax=1280, nlmin=8, nlmax=40)
    meshes = [tbs.ragged_mesh(s) for s in shapes]
    assert sum(mm.nelem for mm in meshes) == 384000
    jacs = [tbs.generate_host(1000 + i, 9, mm.nelem, mm.nintface, 12.0) for i, mm in enumerate(meshes)]
    devj = [(dev(j["jacD"]), dev(j["jacO1"]), dev(j["jacO2"]), dev(j["rhs"])) for j in jacs]
    conv = [oracle.lij_run(mm, j, 12, nthreads=NTHREADS) for mm, j in zip(meshes, jacs)]
    sample = (0, 21, 42, 63)
    for prec in (tbs.TBS_FP64, tbs.TBS_FP32):
        print(f"case cfg3 prec {prec}", flush=True)
        plans = [tbs.Plan(ctx, mm, prec) for mm in meshes]
        tbs.factor_plans(ctx, plans, *[[d[k] for d in devj] for k in range(3)])
        b = tbs.Batch(ctx, plans)
        b.autotune(reps=2)
        Us = [zeros(9 * mm.nelem) for mm in meshes]
        hists, ns = b.solve([d[3] for d in devj], Us, 40, rtol=1e-10, sweeps_per_check=2)
        b.close()
        for p in plans:
            p.close()
        worst_h, worst_u = 0.0, 0.0
        for i, (mm, j) in enumerate(zip(meshes, jacs)):
            U = host(Us[i])
            assert hists[i][-1] <= 1e-10 * hists[i][0], i
            if prec == tbs.TBS_FP64:
                ref = oracle.lij_run(mm, j, int(ns[i]), nthreads=NTHREADS)
                ok, worst = cf.history_close(hists[i], ref["hist"])
                assert ok, (i, worst)
                e = relerr(U, ref["U"])
                assert e <= 1e-12, (i, e)
                worst_h = max(worst_h, worst)
            else:
                e = relerr(U, conv[i]["U"])
                assert e <= 1e-9, (i, e)
            worst_u = max(worst_u, e)
            if i in sample:                       # the same system alone: bitwise
                p = tbs.Plan(ctx, mm, prec)
                p.factor(*devj[i][:3])
                Ua = zeros(9 * mm.nelem)
                ha, na = p.solve(devj[i][3], Ua, 40, 1e-10)
                p.close()
                assert na == ns[i] and np.array_equal(ha, hists[i]) and np.array_equal(host(Ua), U), i
        print(f"cfg3 prec {prec}: worst history dev {worst_h:.3e}, worst U relerr {worst_u:.3e}", flush=True)


def _stream(n_req=24, seed=11):
    """Seeded request shapes (NL, NK, seed) and Poisson arrival times (s)."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(3840), np.log(384000), n_req))
    L = rng.integers(64, 961, n_req)
    NL = np.maximum(1, np.rint(n / L)).astype(int)
    shapes = [(int(a), int(b), 7000 + k) for k, (a, b) in enumerate(zip(NL, L))]
    shapes[5] = (400, 640, 7005)                      # 256,000 elements
    arrivals = np.cumsum(rng.exponential(0.004, n_req))
    return shapes, arrivals


def _serve(tbs, ctx, shapes, arrivals, prec):
    srv = tbs.Server(ctx, precision=prec, rtol=1e-10, max_sweeps=60, sweeps_per_step=2)
    t0 = srv.info()["now"]
    rid_of, done, k = {}, {}, 0
    for _ in range(200000):
        now = srv.info()["now"] - t0
        while k < len(shapes) and arrivals[k] <= now:
            NL, NK, seed = shapes[k]
            rid_of[srv.submit(NL, NK, seed, 12.0)] = k
            k += 1
        srv.step()
        for rid in srv.poll():
            done[rid_of[rid]] = srv.result(rid)
            srv.release(rid)
        if len(done) == len(shapes):
            break
    srv.close()
    assert len(done) == len(shapes)
    return done


def test_cfg5_stream_against_oracle(tbs, ctx, oracle):
    shapes, arrivals = _stream()
    assert max(a * b for a, b, _ in shapes) > 200000
    print(f"case cfg5 stream of {len(shapes)} requests, fp32 factors", flush=True)
    done = _serve(tbs, ctx, shapes, arrivals, tbs.TBS_FP32)
    systems = {}
    worst = 0.0
    for k, (NL, NK, seed) in enumerate(shapes):
        m = tbs.structured_mesh(NL, NK)
        j = tbs.generate_host(seed, 9, m.nelem, m.nintface, 12.0)
        systems[k] = (m, j)
        r = done[k]
        assert r["status"] == 1 and r["hist"][-1] <= 1e-10 * r["hist"][0], (NL, NK)
        ref = oracle.lij_run(m, j, 12, nthreads=NTHREADS)
        e = relerr(r["U"], ref["U"])
        assert e <= 1e-9, (NL, NK, e)
        worst = max(worst, e)
    print(f"cfg5 fp32 server: worst U relerr vs converged oracle {worst:.3e}", flush=True)
    # FP64 server on every third request of the stream: the oracle's same sweeps
    sub = list(range(0, len(shapes), 3)) + [5]
    print(f"case cfg5 fp64 server, {len(sub)} requests", flush=True)
    done64 = _serve(tbs, ctx, [shapes[k] for k in sub], arrivals[: len(sub)], tbs.TBS_FP64)
    for i, k in enumerate(sub):
        m, j = systems[k]
        r = done64[i]
        assert r["status"] == 1
        ref = oracle.lij_run(m, j, r["nsweeps"], nthreads=NTHREADS)
        ok, w = cf.history_close(r["hist"], ref["hist"])
        assert ok, (shapes[k], w)
        assert relerr(r["U"], ref["U"]) <= 1e-12, shapes[k]
