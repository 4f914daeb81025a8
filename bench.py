#!/usr/bin/env python3
"""LIJ sweep benchmark on one MI355X.

Workload: the 1,536,000-element system (1,200 lines x 1,280 elements, 9x9 FP64
blocks, diagonal shift 12, seed 4) from the seeded structured-grid generator,
generated directly in HBM.  One step = one LIJ sweep (Thomas solve + U update,
residual R = rhs + J U, its norm), factorisation excluded; the timed region is
one replay of the K sweeps captured in a hipGraph, as the library's solve runs
them.  A separate eager pass of K sweeps with HIP events around every kernel on
the library's stream gives the per-kernel durations of the roofline (each
kernel launched with hipExtLaunchKernelGGL, whose start/stop events are taken at
the kernel's own start and end).  Prints ONE JSON line.

Also measured (outside the timed region, reported as extra fields): the same
sweep with FP32 factors, hipGraph launch throughput, time to relative residual
1e-10 including factorisation for FP64 and FP32 factors, and the oracle's CPU
throughput on a bounded slice of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, published


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nl", type=int, default=1200)
    ap.add_argument("--nk", type=int, default=1280)
    ap.add_argument("--shift", type=float, default=12.0)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="only the FP64 timed region")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="before timing, copy jacD once (copy_kernel, 81*8*nelem bytes read and written): "
                         "the known-size launch that rocprofv3 FETCH_SIZE/WRITE_SIZE counts are scaled by")
    return ap.parse_args()


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t[workload][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    flags = ""
    try:
        for line in open(os.path.join(ROOT, "oracle", "Makefile")):
            if line.startswith("CXXFLAGS"):
                flags = line.split("=", 1)[1].strip()
    except OSError:
        pass
    return model, flags


def cpu_baseline(tbs, args, nelem_full, sweeps=10):
    """The oracle (the reference's Serial/OpenMP path restated in C++, OpenMP
    over lines) on a bounded slice of the workload: the first tenth of the
    lines of the same grid and generator, `sweeps` timed sweeps after one
    warm-up, with all host cores (up to 16) and with one thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_py import Oracle
    cores = min(os.cpu_count() or 1, 16)
    nl = max(1, args.nl // 10)
    m = tbs.structured_mesh(nl, args.nk)
    jac = tbs.generate_host(args.seed, 9, m.nelem, m.nintface, args.shift)
    o = Oracle()
    scale = m.nelem / nelem_full
    res = {}
    for threads in (cores, 1):
        o.lij_run(m, jac, 1, nthreads=threads)                 # warm-up
        out = o.lij_run(m, jac, sweeps, nthreads=threads)
        res[threads] = (out["t_sweeps"] / sweeps, out["t_factor"])
    model, flags = cpu_info()
    t_all, f_all = res[cores]
    t_one, f_one = res[1]
    return {"value": scale / t_all, "unit": "sweeps/s", "cores": cores, "kind": "port",
            "sample": f"{nl} x {args.nk} slice ({m.nelem} elements, {scale:.3f} of the workload), "
                      f"{sweeps} sweeps after 1 warm-up, OpenMP over lines with {cores} threads; "
                      f"sweeps/s scaled by the element ratio",
            "serial_sweeps_per_s": scale / t_one, "t_sweep_sample_s": t_all, "t_sweep_sample_serial_s": t_one,
            "t_factor_sample_s": f_all, "t_factor_sample_serial_s": f_one, "cpu_model": model,
            "nproc": os.cpu_count(), "compiler_flags": flags}


def bench_batched(tbs, ctx, args, seed=3):
    """64 independent systems of uneven lines (96..1280 elements, 8..40 lines
    per system, 384,000 elements in all) advanced by one launch per kernel per
    sweep, FP64 and FP32-factor; Thomas variant chosen by the autotuner."""
    import numpy as np
    shapes = tbs.batch_shapes(seed, nsys=64, total=384000, lmin=96, lmax=1280, nlmin=8, nlmax=40)
    meshes = [tbs.ragged_mesh(s) for s in shapes]
    res = {"workload": f"batch64_384000_seed{seed}", "nsys": len(meshes),
           "nelem": int(sum(m.nelem for m in meshes)), "nline": int(sum(m.nline for m in meshes))}
    bufs = []
    for i, m in enumerate(meshes):
        z = lambda k: ctx.array(size=k, zero=False)
        D, O1, O2, rhs = z(81 * m.nelem), z(81 * m.nintface), z(81 * m.nintface), z(9 * m.nelem)
        ctx.generate_device(1000 + i, m.nelem, m.nintface, args.shift, D, O1, O2, rhs)
        bufs.append((D, O1, O2, rhs))
    ctx.sync()
    for prec, name in ((tbs.TBS_FP64, "fp64"), (tbs.TBS_FP32, "fp32_factor")):
        plans = [tbs.Plan(ctx, m, prec) for m in meshes]
        ctx.sync()
        t0 = time.perf_counter()
        tbs.factor_plans(ctx, plans, *[[b[k] for b in bufs] for k in range(3)])   # one launch for all
        t_factor = time.perf_counter() - t0
        b = tbs.Batch(ctx, plans)
        chosen, us = b.autotune(reps=3)
        model = sum(p.info()["alg_bytes_per_sweep"] for p in plans)
        steps = args.steps
        for p, (_, _, _, rhs) in zip(plans, bufs):
            p.begin(rhs, None)
        b.launch_sweeps(args.warmup, 4000)
        ctx.sync()
        e0, e1 = ctx.event(), ctx.event()
        ctx.record(e0)
        b.launch_sweeps(steps, 4000)
        ctx.record(e1)
        ctx.sync()
        t = ctx.elapsed_ms(e0, e1) * 1e-3
        ctx.destroy_event(e0)
        ctx.destroy_event(e1)
        # time to 1e-10 from U = 0 (factorisation timed above, solve here)
        Us = [ctx.array(size=9 * m.nelem) for m in meshes]
        ctx.sync()
        t1 = time.perf_counter()
        hists, ns = b.solve([x[3] for x in bufs], Us, 100, rtol=1e-10, sweeps_per_check=2)
        ctx.sync()
        t_solve = time.perf_counter() - t1
        res[name] = {"sweeps_per_s": steps / t, "gbps_model": model / (t / steps) / 1e9,
                     "frac_hbm_peak": model / (t / steps) / 1e9 / HBM_PEAK_GBS, "bytes_per_sweep_model": model,
                     "autotune_variant": int(chosen), "factor_s": t_factor, "solve_to_1e-10_s": t_solve,
                     "time_to_1e-10_s": t_factor + t_solve, "max_sweeps_to_1e-10": int(max(ns)),
                     "worst_final_rel_residual": float(max(h[-1] / h[0] for h in hists))}
        b.close()
        for p in plans:
            p.close()
        for u in Us:
            u.close()
    for t in bufs:
        for x in t:
            x.close()
    return res


def bench_gmres(tbs, ctx, m, args, restart=30):
    """Line-preconditioned GMRES(30) against plain LIJ, both to 1e-10 on the
    headline system with one FP64 factorisation (one LIJ sweep per
    preconditioner application)."""
    n, nif = m.nelem, m.nintface
    out = {}
    for shift, sweeps in ((args.shift, 1),):
        z = lambda k: ctx.array(size=k, zero=False)
        D, O1, O2, rhs = z(81 * n), z(81 * nif), z(81 * nif), z(9 * n)
        ctx.generate_device(args.seed, n, nif, shift, D, O1, O2, rhs)
        U = ctx.array(size=9 * n)
        p = tbs.Plan(ctx, m, tbs.TBS_FP64)
        ctx.sync()
        t0 = time.perf_counter()
        p.factor(D, O1, O2)
        ctx.sync()
        t_factor = time.perf_counter() - t0
        res = {"factor_s": t_factor}
        t0 = time.perf_counter()
        try:
            hist, ns = p.solve(rhs, U, 200, 1e-10)
            ctx.sync()
            res["lij"] = {"sweeps": int(ns), "solve_s": time.perf_counter() - t0,
                          "final_rel_residual": float(hist[-1] / hist[0])}
        except tbs.TbsError as e:
            res["lij"] = {"error": str(e)}
        U.zero()
        g = tbs.Gmres(p, restart)
        t0 = time.perf_counter()
        hist, it = g.solve(rhs, U, 300, 1e-10, sweeps)
        ctx.sync()
        t = time.perf_counter() - t0
        st = g.stats()
        res["gmres"] = {"iterations": int(it), "sweeps_per_preconditioner": sweeps, "solve_s": t, "ms_per_iteration": t / max(it, 1) * 1e3,
                        "final_rel_residual": st["true_residual"] / float(hist[0]), "restarts": st["restarts"]}
        out[f"shift{shift:g}"] = res
        g.close()
        p.close()
        for b_ in (D, O1, O2, rhs, U):
            b_.close()
    return out


def bench_serving(tbs, ctx, args, nreq=160, seed=7):
    """Continuous-batching server, FP32 factors / FP64 residual, tolerance
    1e-10.  Request sizes n log-uniform in [3840, 384000], line length L
    uniform in [64, 960], grid max(1, round(n/L)) x L.  Saturated throughput
    from a burst of nreq requests; then Poisson arrivals at 0.5, 0.8 and 0.95
    of it, latency = scheduled arrival -> result on the host."""
    import numpy as np
    rng = np.random.default_rng(seed)

    def draw(k):
        n = np.exp(rng.uniform(np.log(3840), np.log(384000), k))
        L = rng.integers(64, 961, k)
        NL = np.maximum(1, np.rint(n / L)).astype(int)
        return [(int(a), int(b)) for a, b in zip(NL, L)]

    def run(sizes, arrivals):
        srv = tbs.Server(ctx, precision=tbs.TBS_FP32, rtol=1e-10, max_sweeps=100, sweeps_per_step=2)
        t0 = srv.info()["now"]
        pending = list(range(len(sizes)))
        rid_of, lat, nsw, status = {}, [], [], []
        k = 0
        while len(lat) < len(sizes):
            now = srv.info()["now"] - t0
            while k < len(sizes) and arrivals[k] <= now:
                NL, NK = sizes[k]
                rid_of[srv.submit(NL, NK, 5000 + k, args.shift)] = k
                k += 1
            srv.step()
            for rid in srv.poll():
                r = srv.result(rid, want_U=False)
                lat.append(r["t_done"] - t0 - arrivals[rid_of[rid]])
                nsw.append(r["nsweeps"])
                status.append(r["status"])
                srv.release(rid)
        wall = srv.info()["now"] - t0
        stats = srv.stats()
        srv.close()
        run.stats = stats
        return np.array(lat), wall, nsw, status

    sizes = draw(nreq)
    elems = sum(a * b for a, b in sizes)
    lat, wall, nsw, st = run(sizes, np.zeros(nreq))
    sat = nreq / wall
    out = {"workload": f"serve_{nreq}req_seed{seed}", "mean_elements": elems / nreq,
           "saturated_req_per_s": sat, "saturated_elements_per_s": elems / wall,
           "all_converged": bool(all(s == 1 for s in st)), "max_sweeps": int(max(nsw)),
           "saturated_host_phase_s": run.stats, "loads": {}}
    for frac in (0.5, 0.8, 0.95):
        sizes = draw(nreq)
        arr = np.cumsum(rng.exponential(1.0 / (frac * sat), nreq))
        lat, wall, nsw, st = run(sizes, arr)
        out["loads"][f"{frac:g}"] = {"offered_req_per_s": frac * sat, "achieved_req_per_s": nreq / wall,
                                     "p50_ms": float(np.percentile(lat, 50) * 1e3),
                                     "p99_ms": float(np.percentile(lat, 99) * 1e3),
                                     "all_converged": bool(all(s == 1 for s in st)),
                                     "host_phase_s": run.stats}
    return out


def bench_block_sizes(tbs, args, NL=600, NK=640, sweeps=4, nvars=(5, 9, 10, 16, 17, 22, 32)):
    """Sweeps/s and GB/s against the block size on one 384,000-element grid (the
    paper's block-size study, Fig. 8): the reference loop through the drop-in
    kernels (solveDU, addAtoB, copyAtoB, lineRes: src/main.cxx:207-229) for
    every size, the fused sweep where it is built (2..16)."""
    import numpy as np
    out = {"workload": f"grid_{NL}x{NK}_shift{{2nv+2}}", "nelem": NL * NK, "sweeps": sweeps, "by_nvar": {}}
    plan_nv = set()
    import ctypes
    buf = (ctypes.c_int * 64)()
    for i in range(tbs.load().tbs_plan_supported_nvar(buf, 64)):
        plan_nv.add(buf[i])
    for nv in nvars:
        ctx = tbs.Context(0, nv)
        m = tbs.structured_mesh(NL, NK, nv)
        n, nif = m.nelem, m.nintface
        z = lambda k: ctx.array(size=k, zero=True)
        D, O1, O2, rhs = z(nv * nv * n), z(nv * nv * nif), z(nv * nv * nif), z(nv * n)
        ctx.generate_device(11, n, nif, 2.0 * nv + 2, D, O1, O2, rhs)
        d = m.to_device(ctx)
        DLU, DinvC, Am, U, dU, R = (z(nv * nv * n), z(nv * nv * n), z(nv * nv * n), z(nv * n), z(nv * n), z(nv * n))
        ctx.assembleTriBlocks(d, O1, O2, Am)
        ctx.copyAtoBjac(n, D, DLU)
        ctx.lineLU(n, nif, d, DLU, O1, O2, DinvC)
        ctx.copyAtoB(n, rhs, R)

        def sweep():
            ctx.solveDU(n, nif, d, DLU, DinvC, Am, dU, R)
            ctx.addAtoB(n, dU, U)
            ctx.copyAtoB(n, rhs, R)
            ctx.lineRes(n, nif, d, D, O1, O2, U, R)
        sweep()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(sweeps):
            sweep()
        ctx.sync()
        t_ref = (time.perf_counter() - t0) / sweeps
        ref_bytes = tbs.reference_model_bytes(m)
        row = {"dropin_sweeps_per_s": 1.0 / t_ref, "dropin_gbps_reference_model": ref_bytes / t_ref / 1e9}
        if nv in plan_nv:
            plan = tbs.Plan(ctx, m, tbs.TBS_FP64)
            plan.factor(D, O1, O2)
            plan.begin(rhs, None)
            plan.launch_thomas()
            plan.launch_residual()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(sweeps):
                plan.launch_thomas()
                plan.launch_residual()
            ctx.sync()
            t_f = (time.perf_counter() - t0) / sweeps
            row["fused_sweeps_per_s"] = 1.0 / t_f
            row["fused_gbps_model"] = plan.info()["alg_bytes_per_sweep"] / t_f / 1e9
            plan.close()
        out["by_nvar"][str(nv)] = row
        for x in (D, O1, O2, rhs, DLU, DinvC, Am, U, dU, R, *d.values()):
            x.close()
        ctx.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is designed "
                                   "(DESIGN.md) but not built", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import triblocksolve_amd as tbs

    ctx = tbs.Context(0, 9)
    m = tbs.structured_mesh(args.nl, args.nk)
    n, nif = m.nelem, m.nintface
    z = lambda k: ctx.array(size=k, zero=False)
    D, O1, O2, rhs = z(81 * n), z(81 * nif), z(81 * nif), z(9 * n)
    ctx.generate_device(args.seed, n, nif, args.shift, D, O1, O2, rhs)
    ctx.sync()
    if args.calibrate_copy:
        Dc = z(81 * n)
        ctx.copyAtoBjac(n, D, Dc)
        ctx.sync()
        Dc.close()

    def timed_sweeps(prec, steps, warmup):
        plan = tbs.Plan(ctx, m, prec)
        t0 = time.perf_counter()
        plan.factor(D, O1, O2)
        ctx.sync()
        t_factor = time.perf_counter() - t0
        info = plan.info()
        tuned = plan.autotune(reps=3)
        plan.begin(rhs, None)
        for _ in range(warmup):
            plan.launch_thomas()
            plan.launch_residual()
        ctx.sync()
        # each kernel launched with its own start/stop event pair, recorded
        # by the kernel's dispatch at its start and end
        parts = (tbs.TBS_STEP_THOMAS, tbs.TBS_STEP_RESIDUAL, tbs.TBS_STEP_NORM)
        ev = [[(ctx.event(), ctx.event()) for _ in parts] for _ in range(steps)]
        ctx.sync()
        t0 = time.perf_counter()
        for s in range(steps):
            for k, step in enumerate(parts):
                plan.launch_kernel(step, *ev[s][k])
        ctx.sync()
        wall = time.perf_counter() - t0
        dur = np.array([[ctx.elapsed_ms(*ev[s][k]) for k in range(3)] for s in range(steps)]) * 1e-3
        th, rs, nm = dur[:, 0], dur[:, 1], dur[:, 2]
        for trio in ev:
            for pair in trio:
                for e in pair:
                    ctx.destroy_event(e)
        # the production path (tbs_plan_solve): the K sweeps captured in one
        # hipGraph; the first launch captures and is untimed
        ctx.sync()
        plan.launch_sweeps(steps, use_graph=True)
        ctx.sync()
        t1 = time.perf_counter()
        plan.launch_sweeps(steps, use_graph=True)
        ctx.sync()
        t_graph = time.perf_counter() - t1
        plan.close()
        return dict(wall=wall, thomas=th, residual=rs, norm=nm, info=info, t_factor=t_factor, t_graph=t_graph,
                    tuned=tuned)

    def time_to_tol(prec, rtol=1e-10):
        plan = tbs.Plan(ctx, m, prec)
        U = ctx.array(size=9 * n)
        ctx.sync()
        t0 = time.perf_counter()
        plan.factor(D, O1, O2)
        ctx.sync()
        t1 = time.perf_counter()
        hist, ns = plan.solve(rhs, U, 100, rtol)
        ctx.sync()
        t2 = time.perf_counter()
        plan.close()
        return dict(factor_s=t1 - t0, solve_s=t2 - t1, total_s=t2 - t0, sweeps=int(ns),
                    final_rel_residual=float(hist[-1] / hist[0]), U=U)

    r64 = timed_sweeps(tbs.TBS_FP64, args.steps, args.warmup)
    info = r64["info"]
    # value: the K sweeps as the library runs them (one graph replay between
    # two synchronisations); the event-instrumented eager pass gives the
    # per-kernel durations of the roofline and is reported beside it
    ms_per_step = r64["t_graph"] / args.steps * 1e3
    sweeps_per_s = args.steps / r64["t_graph"]
    model = info["alg_bytes_per_sweep"]
    gbs = model / (r64["t_graph"] / args.steps) / 1e9
    t_th, t_rs = float(r64["thomas"].mean()), float(r64["residual"].mean())
    kern = {
        "thomas": {"avg_ms": t_th * 1e3, "alg_bytes": info["thomas_bytes"],
                   "GBps": info["thomas_bytes"] / t_th / 1e9},
        "residual": {"avg_ms": t_rs * 1e3, "alg_bytes": info["residual_bytes"],
                     "GBps": info["residual_bytes"] / t_rs / 1e9},
        "norm": {"avg_ms": float(r64["norm"].mean()) * 1e3},
    }
    dom = "residual" if t_rs >= t_th else "thomas"
    roof = {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["GBps"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": kern[dom]["GBps"] / HBM_PEAK_GBS,
            "traffic": measured_traffic(f"lij_{args.nl}x{args.nk}", dom)}
    out = {
        "metric": "LIJ sweeps/s at 1.5M elem 9x9 (FP64)",
        "value": sweeps_per_s,
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # BASELINE.md's V100 figure is derived (Table 5 time / 5 sweeps), not published as sweeps/s
        "v100_derived_sweeps_per_s": 65.16,
        "dtype": "f64",
        "data": "synthetic (seeded structured-grid generator, generated in HBM)",
        "config": {"workload": f"lij_{args.nl}x{args.nk}_nvar9_shift{args.shift:g}_seed{args.seed}",
                   "nelem": n, "nline": args.nl, "line_len": args.nk, "nintface": nif,
                   "precision": "fp64", "parallelism": "1 GPU"},
        "gbps_model": gbs,
        "frac_hbm_peak": gbs / HBM_PEAK_GBS,
        "bytes_per_sweep_model": model,
        "bytes_per_sweep_reference_model": tbs.reference_model_bytes(m),
        "gbps_reference_model": tbs.reference_model_bytes(m) / (r64["t_graph"] / args.steps) / 1e9,
        "kernels": kern,
        "roofline": roof,
        "eager_sweeps_per_s": args.steps / r64["wall"],
        "timed_path": "hipGraph of the K sweeps (tbs_plan_launch_sweeps); eager_sweeps_per_s: the same sweeps "
                      "launched one kernel at a time with timing events between them",
        "factor_s": r64["t_factor"],
        "autotune": {"chosen_variant": r64["tuned"][0], "thomas_us_by_variant": r64["tuned"][1]},
    }
    if not args.no_extras:
        r32 = timed_sweeps(tbs.TBS_FP32, args.steps, args.warmup)
        i32 = r32["info"]
        out["fp32_factor"] = {
            "sweeps_per_s": args.steps / r32["t_graph"],
            "gbps_model": i32["alg_bytes_per_sweep"] / (r32["t_graph"] / args.steps) / 1e9,
            "thomas_ms": float(r32["thomas"].mean()) * 1e3,
            "residual_ms": float(r32["residual"].mean()) * 1e3,
            "eager_sweeps_per_s": args.steps / r32["wall"],
        }
        a = time_to_tol(tbs.TBS_FP64)
        b = time_to_tol(tbs.TBS_FP32)
        ua, ub = a["U"].numpy(), b["U"].numpy()
        err = float(np.linalg.norm(ub - ua) / np.linalg.norm(ua))
        for d in (a, b):
            d.pop("U")
        out["time_to_1e-10"] = {"fp64": a, "fp32_factor": b, "fp32_vs_fp64_U_relerr": err}
        out["gmres"] = bench_gmres(tbs, ctx, m, args)
        out["batched"] = bench_batched(tbs, ctx, args)
        out["serving"] = bench_serving(tbs, ctx, args)
        out["block_sizes"] = bench_block_sizes(tbs, args)
    if not args.no_cpu:
        cb = cpu_baseline(tbs, args, n)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline_detail"] = {k: v for k, v in cb.items() if k not in out["cpu_baseline"]}
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
