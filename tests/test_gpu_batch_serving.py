"""GPU: batched uneven lines (many systems per launch) and the continuous-
batching server.  Bar: every system's result is bitwise the one it gets when
solved alone, converges to 1e-10, and its history matches the oracle
(normalised history within 1e-12 * max(h, 1e-4) for FP64)."""
import numpy as np
import pytest

import closed_forms as cf
from gpu_util import dev, host, zeros

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(tbs):
    c = tbs.Context(0, 9)
    yield c
    c.close()


def _solve_alone(tbs, ctx, m, j, prec, iters, rtol):
    p = tbs.Plan(ctx, m, prec)
    p.factor(dev(j["jacD"]), dev(j["jacO1"]), dev(j["jacO2"]))
    U = zeros(9 * m.nelem)
    hist, ns = p.solve(dev(j["rhs"]), U, iters, rtol)
    p.close()
    return hist, ns, host(U)


@pytest.mark.parametrize("one_launch", [False, True], ids=["factor_each", "factor_plans"])
@pytest.mark.parametrize("prec", [0, 1])
def test_uneven_batch_autotuned_graph(tbs, ctx, oracle, prec, one_launch):
    """The systems of a batch, factored one by one or all in one launch
    (tbs_plans_factor), solve bitwise as each does alone (factored alone)."""
    print(f"case uneven batch prec {prec} one_launch {one_launch}", flush=True)
    shapes = tbs.batch_shapes(5, nsys=8, total=30000, lmin=96, lmax=1280, nlmin=3, nlmax=9)
    meshes = [tbs.ragged_mesh(s) for s in shapes]
    jacs = [tbs.generate_host(100 + i, 9, m.nelem, m.nintface, 20.0) for i, m in enumerate(meshes)]
    plans = [tbs.Plan(ctx, m, prec) for m in meshes]
    devj = [(dev(j["jacD"]), dev(j["jacO1"]), dev(j["jacO2"])) for j in jacs]
    if one_launch:
        tbs.factor_plans(ctx, plans, *zip(*devj))
    else:
        for p, (D, O1, O2) in zip(plans, devj):
            p.factor(D, O1, O2)
    b = tbs.Batch(ctx, plans)
    chosen, us = b.autotune(reps=2)
    Us = [zeros(9 * m.nelem) for m in meshes]
    hists, ns = b.solve([dev(j["rhs"]) for j in jacs], Us, 40, rtol=1e-10, sweeps_per_check=4)
    b.close()
    for p in plans:
        p.close()
    for i, (m, j) in enumerate(zip(meshes, jacs)):
        h, n, U = _solve_alone(tbs, ctx, m, j, prec, 40, 1e-10)
        assert ns[i] == n and np.array_equal(hists[i], h) and np.array_equal(host(Us[i]), U), i
        assert h[-1] <= 1e-10 * h[0]
        if prec == 0:
            ref = oracle.lij_run(m, j, n)
            assert cf.history_close(h, ref["hist"])[0]
            assert np.linalg.norm(U - ref["U"]) <= 1e-12 * np.linalg.norm(ref["U"])


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_tuning_tuples_bitwise(tbs, ctx, prec):
    """Tuples the autotuner can pick: two lines per Thomas workgroup (lines of
    different systems in one workgroup, some systems converged while the
    others sweep on), residual workgroups of 1 and 8 waves, eager sweeps on one
    stream and as two halves on two streams, and the graph-captured solve: every
    system's history and U stay bitwise those of its solo solve."""
    print(f"case batch tuning tuples prec {prec}", flush=True)
    shapes = tbs.batch_shapes(6, nsys=5, total=12000, lmin=64, lmax=700, nlmin=3, nlmax=7)
    meshes = [tbs.ragged_mesh(s_) for s_ in shapes]
    # different diagonal shifts: the systems converge after different sweeps
    jacs = [tbs.generate_host(200 + i, 9, m.nelem, m.nintface, [12.0, 30.0, 10.0, 20.0, 50.0][i])
            for i, m in enumerate(meshes)]
    alone = [_solve_alone(tbs, ctx, m, j, prec, 60, 1e-10) for m, j in zip(meshes, jacs)]
    plans = [tbs.Plan(ctx, m, prec) for m in meshes]
    tbs.factor_plans(ctx, plans, *zip(*[(dev(j["jacD"]), dev(j["jacO1"]), dev(j["jacO2"])) for j in jacs]))
    rhs = [dev(j["rhs"]) for j in jacs]
    b = tbs.Batch(ctx, plans)
    codes = plans[0].variants()
    lpw2 = next(v for v, c in enumerate(codes) if c >= 10000)
    rw = {w: r for r, w in enumerate(b.tuning()["residual_waves_per_wg_by_variant"])}
    tuples = [dict(thomas_variant=lpw2, residual_variant=rw[8], two_stream=1),
              dict(thomas_variant=lpw2, residual_variant=rw[1], two_stream=0),
              dict(thomas_variant=0, residual_variant=rw[8], two_stream=1)]
    for t in tuples:
        b.set_tuning(**t)
        got = b.tuning()
        assert all(got[k] == v for k, v in t.items()), (t, got)
        assert got["thomas_lines_per_wg"] == (2 if t["thomas_variant"] == lpw2 else 1)
        # eager sweeps (two halves on two streams when two_stream is set)
        for p, r_ in zip(plans, rhs):
            p.begin(r_, None)
        for _ in range(40):
            b.launch_sweeps(3, 60, 1e-10)
            if not b.states()[:, 1].any():
                break
        for i, (m, p) in enumerate(zip(meshes, plans)):
            U = zeros(9 * m.nelem)
            p.end(U)
            h0, n0, U0 = alone[i]
            assert np.array_equal(p.history(), h0) and np.array_equal(host(U), U0), (t, i)
        # the graph-captured solve
        Us = [zeros(9 * m.nelem) for m in meshes]
        hists, ns = b.solve(rhs, Us, 60, rtol=1e-10, sweeps_per_check=3)
        for i in range(len(meshes)):
            h0, n0, U0 = alone[i]
            assert ns[i] == n0 and np.array_equal(hists[i], h0) and np.array_equal(host(Us[i]), U0), (t, i)
    b.close()
    for p in plans:
        p.close()


def test_server_results_equal_solo_solves(tbs, ctx, oracle):
    print("case server mix", flush=True)
    specs = [(3, 64, 11), (40, 96, 12), (7, 960, 13), (120, 320, 14), (1, 64, 15), (60, 64, 16), (25, 500, 17)]
    srv = tbs.Server(ctx, precision=tbs.TBS_FP32, rtol=1e-10, max_sweeps=60, sweeps_per_step=2)
    ids = {}
    for k, (NL, NK, seed) in enumerate(specs):
        ids[srv.submit(NL, NK, seed, 12.0)] = (NL, NK, seed)
        if k % 2:
            srv.step()
    done = {}
    for _ in range(2000):
        a, q, m = srv.step()
        for rid in srv.poll():
            done[rid] = srv.result(rid)
            srv.release(rid)
        if len(done) == len(specs):
            break
    assert len(done) == len(specs)
    for rid, (NL, NK, seed) in ids.items():
        r = done[rid]
        assert r["status"] == 1 and r["hist"][-1] <= 1e-10 * r["hist"][0]
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]
        mesh = tbs.structured_mesh(NL, NK)
        jac = tbs.generate_host(seed, 9, mesh.nelem, mesh.nintface, 12.0)
        h, n, U = _solve_alone(tbs, ctx, mesh, jac, tbs.TBS_FP32, 60, 1e-10)
        assert r["nsweeps"] == n and np.array_equal(r["hist"], h), (NL, NK)
        assert np.array_equal(r["U"], U), (NL, NK)
        # and against the FP64 oracle's solution at the round-off floor
        ref = oracle.lij_run(mesh, jac, 12)
        assert np.linalg.norm(r["U"] - ref["U"]) <= 1e-9 * np.linalg.norm(ref["U"]), (NL, NK)
    srv.close()


def test_server_admits_by_footprint(tbs, ctx):
    """A budget that holds about two requests still completes a queue of six,
    in FIFO order of admission, and refuses a request larger than the budget."""
    print("case server budget", flush=True)
    probe = tbs.Server(ctx, precision=tbs.TBS_FP32)
    fp = probe.footprint(40, 200)
    probe.close()
    assert fp > tbs.Plan.footprint(tbs.structured_mesh(40, 200), tbs.TBS_FP32)
    budget = int(2.5 * fp)
    srv = tbs.Server(ctx, hbm_budget=budget, precision=tbs.TBS_FP32, rtol=1e-10, max_sweeps=60, sweeps_per_step=2)
    rids = [srv.submit(40, 200, 30 + i) for i in range(6)]
    with pytest.raises(tbs.TbsError) as ei:
        srv.submit(2000, 500, 1)
    assert ei.value.code == -4
    done, peak = {}, 0
    for _ in range(4000):
        a, q, m = srv.step()
        peak = max(peak, a + m)
        assert srv.info()["used"] <= budget
        for rid in srv.poll():
            done[rid] = srv.result(rid, want_U=False)
        if len(done) == 6:
            break
    assert len(done) == 6 and all(d["status"] == 1 for d in done.values())
    assert peak >= 2
    st = srv.stats()
    assert st["admitted"] == 6 and st["retired"] == 6
    admits = [done[r]["t_admit"] for r in rids]
    assert admits == sorted(admits)
    srv.close()


def test_plan_footprint_is_exact(tbs, ctx):
    for NL, NK, prec in ((7, 300, 0), (40, 64, 1), (3, 1, 0)):
        m = tbs.structured_mesh(NL, NK)
        p = tbs.Plan(ctx, m, prec)
        assert p.info()["device_bytes"] == tbs.Plan.footprint(m, prec)
        p.close()


@pytest.mark.parametrize("prec", [1, 0], ids=["fp32", "fp64"])
def test_batch_membership_changes_in_place(tbs, ctx, prec):
    """Systems leave and join a running batch between sweeps; every system's
    history and U stay bitwise those of its solo solve (FP64 batches run their
    eager sweeps as two halves on two streams)."""
    print(f"case batch set_plans prec {prec}", flush=True)
    specs = [(5, 200, 41), (12, 96, 42), (2, 700, 43), (30, 64, 44)]
    systems = []
    for NL, NK, seed in specs:
        m = tbs.structured_mesh(NL, NK)
        j = tbs.generate_host(seed, 9, m.nelem, m.nintface, 12.0)
        p = tbs.Plan(ctx, m, prec)
        p.factor(dev(j["jacD"]), dev(j["jacO1"]), dev(j["jacO2"]))
        rhs = dev(j["rhs"])
        systems.append((m, j, p, rhs))
    for m, j, p, rhs in systems[:2]:
        p.begin(rhs, None)
    b = tbs.Batch(ctx, [systems[0][2], systems[1][2]])
    b.launch_sweeps(3, 60, 1e-10)
    for m, j, p, rhs in systems[2:]:
        p.begin(rhs, None)
    b.set_plans([systems[1][2], systems[2][2], systems[3][2], systems[0][2]])
    b.launch_sweeps(2, 60, 1e-10)
    b.set_plans([])
    b.launch_sweeps(2, 60, 1e-10)      # no systems: nothing runs
    b.set_plans([s[2] for s in systems])
    for _ in range(40):
        b.launch_sweeps(2, 60, 1e-10)
        if not b.states()[:, 1].any():
            break
    assert not b.states()[:, 1].any()
    for m, j, p, rhs in systems:
        U = zeros(9 * m.nelem)
        p.end(U)
        h = p.history()
        h0, n0, U0 = _solve_alone(tbs, ctx, m, j, prec, 60, 1e-10)
        assert np.array_equal(h, h0) and np.array_equal(host(U), U0)
    b.close()
    for s_ in systems:
        s_[2].close()


def test_server_fp64_mixed_sizes(tbs, ctx, oracle):
    """FP64 server with tiny (one element, one line) and larger requests in one
    batch: every result is bitwise its solo FP64 solve."""
    print("case server fp64 mixed", flush=True)
    specs = [(1, 1, 51), (1, 300, 52), (90, 8, 53), (200, 150, 54), (3, 2, 55)]
    srv = tbs.Server(ctx, precision=tbs.TBS_FP64, rtol=1e-10, max_sweeps=60, sweeps_per_step=3)
    ids = {srv.submit(NL, NK, seed, 12.0): (NL, NK, seed) for NL, NK, seed in specs}
    done = {}
    for _ in range(3000):
        srv.step()
        for rid in srv.poll():
            done[rid] = srv.result(rid)
        if len(done) == len(specs):
            break
    assert len(done) == len(specs)
    for rid, (NL, NK, seed) in ids.items():
        mesh = tbs.structured_mesh(NL, NK)
        jac = tbs.generate_host(seed, 9, mesh.nelem, mesh.nintface, 12.0)
        h, n, U = _solve_alone(tbs, ctx, mesh, jac, tbs.TBS_FP64, 60, 1e-10)
        r = done[rid]
        assert r["status"] == 1 and r["nsweeps"] == n
        assert np.array_equal(r["hist"], h) and np.array_equal(r["U"], U), (NL, NK)
        # the served result against the oracle directly (same sweeps)
        ref = oracle.lij_run(mesh, jac, n)
        assert cf.history_close(r["hist"], ref["hist"])[0], (NL, NK)
        assert np.linalg.norm(r["U"] - ref["U"]) <= 1e-12 * np.linalg.norm(ref["U"]), (NL, NK)
    srv.close()


def test_server_rejects_bad_requests(tbs, ctx):
    srv = tbs.Server(ctx, hbm_budget=1 << 30, precision=tbs.TBS_FP32)
    for NL, NK in ((0, 10), (10, 0), (-1, 5)):
        with pytest.raises(tbs.TbsError):
            srv.submit(NL, NK, 1)
    with pytest.raises(tbs.TbsError) as ei:
        srv.result(12345)
    assert ei.value.code == -1
    srv.close()


def test_plans_factor_rejects_mixed_plans(tbs, ctx):
    """One launch needs one block size and one precision."""
    m = tbs.ragged_mesh([5, 9])
    j = tbs.generate_host(3, 9, m.nelem, m.nintface, 20.0)
    a, b = tbs.Plan(ctx, m, 0), tbs.Plan(ctx, m, 1)
    D, O1, O2 = dev(j["jacD"]), dev(j["jacO1"]), dev(j["jacO2"])
    with pytest.raises(tbs.TbsError):
        tbs.factor_plans(ctx, [a, b], [D, D], [O1, O1], [O2, O2])
    a.close()
    b.close()


def test_server_serves_callers_systems(tbs, ctx, oracle, tmp_path):
    """A caller's system -- read back from the reference's binary files that
    triblock.exe --write-data writes, through the host API's readers -- served
    from host arrays and again from device arrays, next to a ragged system and
    seeded synthetic requests: every result equals the oracle's (FP64 server:
    history within 1e-12 * max(h, 1e-4), U within 1e-12) and the same system
    solved alone, bitwise."""
    import os
    import subprocess
    from conftest import ROOT
    print("case server caller systems", flush=True)
    exe = os.path.join(ROOT, "triblocksolve_amd", "triblock.exe")
    p = subprocess.run([exe, "1", "0", "9", "--synthetic", "30x200", "--seed", "9", "--shift", "12",
                        "--iters", "1", "--write-data"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    m_file = tbs.read_mesh(os.path.join(tmp_path, "gpuline.mesh.data.bin"))
    nvar, nelem, nif, j_file = tbs.read_jacobian(os.path.join(tmp_path, "gpuline.jacobian.data.bin"))
    assert (nvar, nelem, nif) == (9, m_file.nelem, m_file.nintface)
    m_rag = tbs.ragged_mesh([5, 300, 17, 1, 64, 130])
    j_rag = tbs.generate_host(77, 9, m_rag.nelem, m_rag.nintface, 12.0)
    srv = tbs.Server(ctx, precision=tbs.TBS_FP64, rtol=1e-10, max_sweeps=60, sweeps_per_step=2)
    assert srv.footprint_system(m_file) > tbs.Plan.footprint(m_file, tbs.TBS_FP64)
    ids = {}
    ids[srv.submit_system(m_file, j_file["jacD"], j_file["jacO1"], j_file["jacO2"], j_file["rhs"])] = \
        ("file_host", m_file, j_file)
    ids[srv.submit(20, 96, 31, 12.0)] = ("synthetic", None, None)
    devj = [dev(j_file[k]) for k in ("jacD", "jacO1", "jacO2", "rhs")]
    ids[srv.submit_system(m_file, *devj)] = ("file_device", m_file, j_file)
    ids[srv.submit_system(m_rag, j_rag["jacD"], j_rag["jacO1"], j_rag["jacO2"], j_rag["rhs"])] = \
        ("ragged_host", m_rag, j_rag)
    ids[srv.submit(3, 700, 32, 12.0)] = ("synthetic", None, None)
    # a mesh the fused path cannot take is refused at submit
    bad = tbs.ragged_mesh([4, 4])
    bad.lineface[2] = 0                   # not the face joining the line's elements 1 and 2
    with pytest.raises(tbs.TbsError):
        srv.submit_system(bad, *[np.zeros(81 * 8)] * 3, np.zeros(72))
    done = {}
    for _ in range(5000):
        srv.step()
        for rid in srv.poll():
            done[rid] = (srv.result(rid), srv.timeline(rid))
            srv.release(rid)
        if len(done) == len(ids):
            break
    srv.close()
    assert len(done) == len(ids)
    results = {}
    for rid, (name, m, j) in ids.items():
        r, tl = done[rid]
        assert r["status"] == 1 and r["hist"][-1] <= 1e-10 * r["hist"][0], name
        assert np.all(np.diff(tl) >= 0), (name, tl)
        if m is None:
            continue
        ref = oracle.lij_run(m, j, r["nsweeps"])
        ok, worst = cf.history_close(r["hist"], ref["hist"])
        assert ok, (name, worst)
        assert np.linalg.norm(r["U"] - ref["U"]) <= 1e-12 * np.linalg.norm(ref["U"]), name
        h, n, U = _solve_alone(tbs, ctx, m, j, tbs.TBS_FP64, 60, 1e-10)
        assert r["nsweeps"] == n and np.array_equal(r["hist"], h) and np.array_equal(r["U"], U), name
        results[name] = r
    assert np.array_equal(results["file_host"]["U"], results["file_device"]["U"])
