"""GPU: the fused sweep path (plan / batch) against the oracle through the C ABI.

Tolerances (stated per result):
  FP64  history: |h_k/h_0 - h'_k/h'_0| <= 1e-12 * max(h'_k/h'_0, 1e-4)   (h' = oracle)
        solution: ||U - U'|| <= 1e-12 ||U'||
  FP32 factors: converge to h_k/h_0 <= 1e-10 and ||U - U*|| <= 1e-9 ||U*||,
        U* = FP64 oracle solution run to the round-off floor.
"""
import os

import numpy as np
import pytest

import closed_forms as cf
import synth_np
from conftest import system
from gpu_util import cross_lines, dev, host, mctx, ragged_lines, reversed_lines, zeros

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


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


def solve_plan(tbs, ctx, m, jac, iters, rtol=0.0, prec=0, want_R=False, variant=None, rvariant=None):
    plan = tbs.Plan(ctx, m, prec)
    plan.factor(dev(jac["jacD"]), dev(jac["jacO1"]), dev(jac["jacO2"]))
    if variant is not None:
        plan.set_variant(variant)
    if rvariant is not None:
        plan.set_tuning(residual_variant=rvariant)
    U = zeros(m.nvar * m.nelem)
    R = zeros(U.size) if want_R else None
    hist, ns = plan.solve(dev(jac["rhs"]), U, iters, rtol, R_out=R)
    out = dict(hist=hist, ns=ns, U=host(U), R=host(R) if want_R else None)
    plan.close()
    return out


@pytest.mark.parametrize("shift", [20.0, 12.0])
def test_cfg1_fp64_history_and_solution(tbs, ctxs, oracle, shift):
    print(f"case cfg1 60x64 shift {shift}", flush=True)
    m, jac = system(60, 64, 9, seed=1, shift=shift)
    ref = oracle.lij_run(m, jac, 10)
    out = solve_plan(tbs, ctxs(9), m, jac, 10, want_R=True)
    assert out["ns"] == 10
    ok, worst = cf.history_close(out["hist"], ref["hist"])
    print(f"history worst {worst:.3e}; U relerr {relerr(out['U'], ref['U']):.3e}")
    assert ok, worst
    assert relerr(out["U"], ref["U"]) < 1e-12
    # R_final is rhs + J U for the returned U
    R_cf = jac["rhs"] + cf.assemble_J(m, jac) @ out["U"]
    assert np.allclose(out["R"], R_cf, rtol=0, atol=1e-13 * np.abs(jac["rhs"]).max())


VARIANTS = [
    ("reversed_5x13_nv9", lambda: reversed_lines(system(5, 13, 9)[0])),
    ("cross_7x9_nv5", lambda: cross_lines(7, 9, 5)),
    ("ragged_nv9", lambda: ragged_lines([1, 2, 17, 64, 3, 100, 1, 65, 7, 8, 9], 9)),
    ("grid_4x6_nv4", lambda: system(4, 6, 4)[0]),
    ("grid_1x1_nv9", lambda: system(1, 1, 9)[0]),
    ("grid_70x1_nv9", lambda: system(70, 1, 9)[0]),
]


@pytest.mark.parametrize("name,make", VARIANTS, ids=[v[0] for v in VARIANTS])
def test_variants_fp64(tbs, ctxs, oracle, name, make):
    print(f"case {name}", flush=True)
    m = make()
    jac = synth_np.generate(11, m.nvar, m.nelem, m.nintface, 2.0 * m.nvar + 2)
    ref = oracle.lij_run(m, jac, 8)
    out = solve_plan(tbs, ctxs(m.nvar), m, jac, 8)
    ok, worst = cf.history_close(out["hist"], ref["hist"])
    assert ok, worst
    assert relerr(out["U"], ref["U"]) < 1e-12


@pytest.mark.parametrize("nlines", [1100, 1400])
@pytest.mark.parametrize("prec", [0, 1])
def test_factor_in_two_rounds_ragged(tbs, ctxs, oracle, prec, nlines):
    """1,100 and 1,400 ragged lines (seeded lengths 1..90) around the factor
    kernel's resident workgroups for 9x9 blocks on a 256-CU MI355X (1,024
    with FP64 factors, 1,280 with FP32 factors): workgroups that start only
    when others finish, beside the block packing.  FP64 against the
    oracle's same sweeps; FP32 factors to 1e-9 of the FP64 oracle's
    converged solution."""
    rng = np.random.default_rng(77)
    lens = rng.integers(1, 91, size=nlines)
    m = tbs.ragged_mesh(lens, nvar=9)
    print(f"case ragged {nlines} lines, {m.nelem} elements, prec {prec}", flush=True)
    jac = synth_np.generate(78, 9, m.nelem, m.nintface, 12.0)
    if prec == 0:
        ref = oracle.lij_run(m, jac, 6, nthreads=min(16, os.cpu_count() or 1))
        out = solve_plan(tbs, ctxs(9), m, jac, 6)
        ok, worst = cf.history_close(out["hist"], ref["hist"])
        assert ok, worst
        assert relerr(out["U"], ref["U"]) < 1e-12
    else:
        ref = oracle.lij_run(m, jac, 14, nthreads=min(16, os.cpu_count() or 1))
        out = solve_plan(tbs, ctxs(9), m, jac, 40, rtol=1e-10, prec=1)
        assert out["hist"][-1] <= 1e-10 * out["hist"][0]
        assert relerr(out["U"], ref["U"]) < 1e-9


@pytest.mark.parametrize("prec", [0, 1])
def test_thomas_variants_bitwise(tbs, ctxs, oracle, prec):
    """Every compiled Thomas variant (prefetch depth, LDS windows, lines per
    workgroup) and residual workgroup shape does the same arithmetic in the
    same order: history and U are bitwise equal across variants, on ragged
    lines (ends inside a window, lines of one element, lines longer than a
    window group, an odd line count so a two-line workgroup has an empty
    slot), and the FP64 ones match the oracle."""
    nv = 9                                     # the block size with several compiled variants
    m = ragged_lines([1, 2, 17, 64, 3, 100, 1, 65, 7, 8, 9, 130, 5], nv)
    jac = synth_np.generate(40 + nv, nv, m.nelem, m.nintface, 2.0 * nv + 2)
    ctx = ctxs(nv)
    probe = tbs.Plan(ctx, m, prec)
    names = [tbs.variant_name(c) for c in probe.variants()]
    nres = len(probe.tuning()["residual_us"])
    probe.close()
    print(f"case prec {prec}: variants {names}, {nres} residual shapes", flush=True)
    assert len(names) > 1 and nres > 1
    assert any(n.endswith("_l2") for n in names)
    outs = [solve_plan(tbs, ctx, m, jac, 8, prec=prec, variant=v) for v in range(len(names))]
    outs += [solve_plan(tbs, ctx, m, jac, 8, prec=prec, rvariant=r) for r in range(1, nres)]
    names += [f"residual_{r}" for r in range(1, nres)]
    for name, o in zip(names[1:], outs[1:]):
        assert np.array_equal(o["hist"], outs[0]["hist"]), name
        assert np.array_equal(o["U"], outs[0]["U"]), name
    if prec == 0:
        ref = oracle.lij_run(m, jac, 8)
        ok, worst = cf.history_close(outs[0]["hist"], ref["hist"])
        assert ok, worst
        assert relerr(outs[0]["U"], ref["U"]) < 1e-12


@pytest.mark.parametrize("nv", [2, 3, 6, 7, 8, 10, 11, 12, 13, 14, 15, 16] + list(range(17, 33)))
def test_block_sizes(tbs, ctxs, oracle, nv):
    """Every compiled block size of the fused path (17..32: the two-columns-
    per-lane factor kernel): FP64 history and U against the oracle, FP32
    factors converge to the same solution.  Five sweeps: on these small
    systems (207 elements) the residual reaches the round-off floor (~1e-16
    of h_0) by sweep 7-8 for the larger blocks, where two correct
    implementations differ by the floor itself and the 1e-12 * max(h, 1e-4)
    bound is no longer a parity test."""
    print(f"case block size {nv}", flush=True)
    m = ragged_lines([9, 1, 40, 23, 64, 70], nv)
    jac = synth_np.generate(20 + nv, nv, m.nelem, m.nintface, 2.0 * nv + 2)
    ref = oracle.lij_run(m, jac, 5)
    out = solve_plan(tbs, ctxs(nv), m, jac, 5)
    assert out["hist"][-1] > 1e-13 * out["hist"][0]       # still above the floor
    ok, worst = cf.history_close(out["hist"], ref["hist"])
    assert ok, worst
    assert relerr(out["U"], ref["U"]) < 1e-12
    o32 = solve_plan(tbs, ctxs(nv), m, jac, 40, rtol=1e-10, prec=1)
    assert o32["hist"][-1] <= 1e-10 * o32["hist"][0]
    assert relerr(o32["U"], ref["U"]) < 1e-8


def test_fp32_factors_reach_fp64_accuracy(tbs, ctxs, oracle):
    """FP32 factors and Thomas sweeps, FP64 residual/update/norm: same 1e-10 target."""
    m, jac = system(60, 64, 9, seed=4, shift=12.0)
    ref = oracle.lij_run(m, jac, 12)
    o64 = solve_plan(tbs, ctxs(9), m, jac, 50, rtol=1e-10, prec=0)
    o32 = solve_plan(tbs, ctxs(9), m, jac, 50, rtol=1e-10, prec=1)
    print(f"sweeps fp64 {o64['ns']} fp32 {o32['ns']}; fp32 U err vs oracle {relerr(o32['U'], ref['U']):.3e}")
    assert o64["hist"][-1] <= 1e-10 * o64["hist"][0]
    assert o32["hist"][-1] <= 1e-10 * o32["hist"][0]
    assert o32["ns"] <= o64["ns"] + 2
    assert relerr(o32["U"], ref["U"]) < 1e-9


@pytest.mark.parametrize("nv", [2, 3, 5, 9, 16, 17, 21, 22, 24, 25, 32])
def test_one_sweep_is_the_line_solve(tbs, ctxs, nv):
    """From U = 0 one sweep is U = -M^-1 rhs exactly (M: the lines' block
    tridiagonal, assembled and solved by scipy), whatever the line lengths:
    pins the two-ended factorisation and Thomas pass, whose split point,
    window padding and empty halves all depend on the length (window of
    64 // nv steps).  Lengths around 1, 2, 3 windows and their halves, both
    line orientations; FP64 factors 1e-13, FP32 factors 1e-5."""
    import scipy.sparse.linalg as spla
    spw = 64 // nv
    lens = sorted({1, 2, 3, spw - 1, spw, spw + 1, 2 * spw, 2 * spw + 1, 3 * spw - 1, 3 * spw + 2,
                   4 * spw, 5 * spw + 3, 7 * spw + 1, 13 * spw, 97, 130} - {0})
    print(f"case one sweep nv {nv} lengths {lens}", flush=True)
    base = ragged_lines(lens, nv)
    for m in (base, reversed_lines(base)):
        jac = synth_np.generate(7, nv, m.nelem, m.nintface, 2.0 * nv + 2.0)
        ref = -spla.splu(cf.assemble_M(m, jac).tocsc()).solve(jac["rhs"])
        for prec, tol in ((0, 1e-13), (1, 1e-5)):
            out = solve_plan(tbs, ctxs(nv), m, jac, 1, prec=prec)
            assert out["ns"] == 1
            assert relerr(out["U"], ref) < tol, (prec, relerr(out["U"], ref))


def test_second_solve_equals_first(tbs, ctxs):
    """A plan's first solve launches its sweeps eagerly, later ones replay
    captured graphs: the same history and U, bitwise."""
    m, jac = system(30, 40, 9, seed=5)
    plan = tbs.Plan(ctxs(9), m, 0)
    plan.factor(dev(jac["jacD"]), dev(jac["jacO1"]), dev(jac["jacO2"]))
    out = []
    for _ in range(3):
        U = zeros(9 * m.nelem)
        hist, ns = plan.solve(dev(jac["rhs"]), U, 30, 1e-12)
        out.append((hist, ns, host(U)))
    plan.close()
    for hist, ns, U in out[1:]:
        assert ns == out[0][1] and np.array_equal(hist, out[0][0]) and np.array_equal(U, out[0][2])


def test_graph_replay_equals_eager(tbs, ctxs):
    """Several replays of a captured K-sweep graph, with eager work between them,
    give exactly the eager launches' history and solution."""
    m, jac = system(30, 40, 9, seed=2)
    ctx = ctxs(9)
    res = {}
    steps = (tbs.TBS_STEP_THOMAS, tbs.TBS_STEP_RESIDUAL, tbs.TBS_STEP_NORM)
    for mode in ("eager", "graph", "kernels"):
        plan = tbs.Plan(ctx, m)
        plan.factor(dev(jac["jacD"]), dev(jac["jacO1"]), dev(jac["jacO2"]))
        rhs = dev(jac["rhs"])
        plan.begin(rhs, None)
        for _ in range(3):
            if mode == "kernels":   # one kernel at a time, each timed by its own events
                for _ in range(2):
                    for step in steps:
                        e0, e1 = ctx.event(), ctx.event()
                        plan.launch_kernel(step, e0, e1)
                        ctx.sync()
                        assert ctx.elapsed_ms(e0, e1) > 0.0
                        ctx.destroy_event(e0)
                        ctx.destroy_event(e1)
            else:
                plan.launch_sweeps(2, use_graph=(mode == "graph"))
            ctx.sync()
            zeros(1000).numpy().sum()   # unrelated work and a device-wide sync
        U = zeros(9 * m.nelem)
        plan.end(U)
        ctx.sync()
        res[mode] = (plan.history(), host(U))
        plan.close()
    for mode in ("graph", "kernels"):
        assert np.array_equal(res["eager"][0], res[mode][0]), mode
        assert np.array_equal(res["eager"][1], res[mode][1]), mode


def test_deterministic(tbs, ctxs):
    m, jac = system(40, 50, 9, seed=3)
    a = solve_plan(tbs, ctxs(9), m, jac, 6)
    b = solve_plan(tbs, ctxs(9), m, jac, 6)
    assert np.array_equal(a["U"], b["U"]) and np.array_equal(a["hist"], b["hist"])


def test_batch_equals_alone_and_oracle(tbs, ctxs, oracle):
    """Systems of uneven lines advanced by one launch per kernel: each result is
    bitwise the one it gets alone, and matches the oracle."""
    ctx = ctxs(9)
    specs = [system(20, 30, 9, seed=5)[0], reversed_lines(system(6, 17, 9)[0]),
             ragged_lines([5, 96, 7, 130, 64], 9), system(3, 200, 9)[0]]
    jacs = [synth_np.generate(21 + i, 9, m.nelem, m.nintface, 20.0 - 2 * i) for i, m in enumerate(specs)]
    plans = []
    for m, j in zip(specs, jacs):
        p = tbs.Plan(ctx, m)
        p.factor(dev(j["jacD"]), dev(j["jacO1"]), dev(j["jacO2"]))
        plans.append(p)
    batch = tbs.Batch(ctx, plans)
    Us = [zeros(9 * m.nelem) for m in specs]
    hists, ns = batch.solve([dev(j["rhs"]) for j in jacs], Us, 40, rtol=1e-10, sweeps_per_check=3)
    batch.close()
    for i, (m, j) in enumerate(zip(specs, jacs)):
        alone = solve_plan(tbs, ctx, m, j, 40, rtol=1e-10)
        assert ns[i] == alone["ns"]
        assert np.array_equal(hists[i], alone["hist"]), i
        assert np.array_equal(host(Us[i]), alone["U"]), i
        ref = oracle.lij_run(m, j, int(ns[i]))
        assert cf.history_close(hists[i], ref["hist"])[0]
        assert hists[i][-1] <= 1e-10 * hists[i][0]
    for p in plans:
        p.close()


def test_bad_inputs_raise(tbs, ctxs):
    m, jac = system(4, 5, 9)
    bad = type(m)(m.nvar, m.epoint, m.ef, m.fc, m.linesize, m.linepoint, np.zeros_like(m.lines), m.lineface)
    with pytest.raises(tbs.TbsError):
        tbs.Plan(ctxs(9), bad)
    p = tbs.Plan(ctxs(9), m)
    with pytest.raises(tbs.TbsError) as ei:
        p.solve(dev(jac["rhs"]), zeros(9 * m.nelem), 3)
    assert ei.value.code == -2
    e = ctxs(9).event()
    for step, ev in ((7, (None, None)), (tbs.TBS_STEP_NORM, (e, None))):   # rejected before any launch
        with pytest.raises(tbs.TbsError) as ei:
            p.launch_kernel(step, *ev)
        assert ei.value.code == -1
    ctxs(9).destroy_event(e)
    p.close()
    with pytest.raises(tbs.TbsError):
        tbs.Context(0, 33)


def test_divergence_is_reported(tbs, ctxs):
    """Off-line blocks scaled so the iteration diverges: the solve stops with TBS_ERR_NUMERIC."""
    m, jac = system(10, 12, 9, seed=6, shift=0.5)
    for k in ("jacO1", "jacO2"):
        jac[k] = jac[k] * 40.0
    plan = tbs.Plan(ctxs(9), m)
    plan.factor(dev(jac["jacD"]), dev(jac["jacO1"]), dev(jac["jacO2"]))
    with pytest.raises(tbs.TbsError) as ei:
        plan.solve(dev(jac["rhs"]), zeros(9 * m.nelem), 200, 1e-10)
    assert ei.value.code == -6
    plan.close()
