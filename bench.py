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

Also measured (outside the timed region, reported as extra fields): the
device's achievable streaming bandwidth, the errors against the FP64 oracle,
the same sweep with FP32 factors, time to relative residual 1e-10 including
factorisation for FP64 and FP32 factors, the reference loop on the drop-in
kernels, GMRES, the batched and served workloads, block sizes, and the
oracle's CPU throughput on the same workload.
"""
import argparse
import json
import os
import sys
import time
import traceback

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
    ap.add_argument("--no-oracle", action="store_true", help="skip the errors against the FP64 oracle")
    ap.add_argument("--section", choices=["batched", "serving"], default=None,
                    help="run only this extra section (cfg3 batch or cfg5 serving, without the oracle) and "
                         "print its JSON; for profiling one workload")
    ap.add_argument("--cpu-baseline-only", action="store_true",
                    help="(internal) run only the CPU baseline leg and print its JSON; no GPU use")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="before timing, copy jacD once (copy_kernel, 81*8*nelem bytes read and written): "
                         "the known-size launch that rocprofv3 FETCH_SIZE/WRITE_SIZE counts are scaled by")
    return ap.parse_args()


def log(msg):
    """Progress of the run on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def host_cores():
    """Host cores this process may use, at most 16 (the GPU box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline_child(args):
    """Runs in a child process started by cpu_baseline() with OMP_PROC_BIND=close
    set before the OpenMP runtime loads: the oracle (the reference's
    Serial/OpenMP path restated in C++, OpenMP over lines) on the whole
    workload, 11 sweeps each with all cores and with one thread; the first
    sweep is the warm-up, the median of the other 10 is the figure.  No norms
    are taken inside the timed sweeps (the reference computes none)."""
    import numpy as np
    import triblocksolve_amd as tbs
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_py import Oracle
    cores = host_cores()
    m = tbs.structured_mesh(args.nl, args.nk)
    jac = tbs.generate_host(args.seed, 9, m.nelem, m.nintface, args.shift)
    o = Oracle()
    res = {}
    for threads in (cores, 1):
        out = o.lij_run(m, jac, 11, nthreads=threads, norms=False)
        t = np.asarray(out["t_each"][1:11])
        res[threads] = (float(np.median(t)), float(out["t_factor"]), [float(x) for x in t])
    model, flags = cpu_info()
    t_all, f_all, each = res[cores]
    t_one, f_one, _ = res[1]
    print(json.dumps({
        "value": 1.0 / t_all, "unit": "sweeps/s", "cores": cores, "kind": "port",
        "sample": f"the whole {args.nl} x {args.nk} workload ({m.nelem} elements): median of 10 sweeps after "
                  f"1 warm-up sweep, OpenMP over lines with {cores} threads, OMP_PROC_BIND=close",
        "serial_sweeps_per_s": 1.0 / t_one, "t_sweep_median_s": t_all, "t_sweep_median_serial_s": t_one,
        "t_sweeps_s": each, "t_factor_s": f_all, "t_factor_serial_s": f_one, "cpu_model": model,
        "nproc": os.cpu_count(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
        "compiler_flags": flags}))


def cpu_baseline(args):
    """The CPU leg in a child process (no GPU use there), so that OMP_PROC_BIND
    and OMP_PLACES reach the OpenMP runtime before it starts."""
    import subprocess
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores", OMP_NUM_THREADS=str(host_cores()))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--nl", str(args.nl),
           "--nk", str(args.nk), "--shift", str(args.shift), "--seed", str(args.seed)]
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError(f"CPU baseline failed ({out.returncode}): {out.stderr[-2000:]}")
    return json.loads(out.stdout.strip().splitlines()[-1])


def relerr(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def hist_dev(h, ref, floor=1e-4):
    """Largest |h_k/h_0 - h'_k/h'_0| / max(h'_k/h'_0, floor): the parity measure of
    the residual history against the oracle's (1e-12 is the bound the tests use)."""
    import numpy as np
    n = min(len(h), len(ref))
    a = np.asarray(h[:n]) / h[0]
    b = np.asarray(ref[:n]) / ref[0]
    return float(np.max(np.abs(a - b) / np.maximum(b, floor)))


def tuning_tuple(t):
    """The autotuner's choice: Thomas prefetch windows, LDS windows per group
    and lines per workgroup, residual waves per workgroup, two-stream split."""
    return {k: t[k] for k in ("thomas_prefetch", "thomas_lds_windows", "thomas_lines_per_wg",
                              "residual_waves_per_wg", "two_stream")}


def residual_by_rw(t):
    """Residual launch time (us) per workgroup shape, from the last autotune."""
    return {str(w): us for w, us in zip(t["residual_waves_per_wg_by_variant"], t["residual_us"])}


def bench_batched(tbs, ctx, args, seed=3, oracle=None):
    """64 independent systems of uneven lines (96..1280 elements, 8..40 lines
    per system, 384,000 elements in all) advanced by one launch per kernel per
    sweep, FP64 and FP32-factor; Thomas variant chosen by the autotuner.  With
    the oracle: every system's U against the FP64 oracle (FP64 batch: the same
    number of sweeps; FP32 factors: the oracle run to the round-off floor)."""
    import numpy as np
    shapes = tbs.batch_shapes(seed, nsys=64, total=384000, lmin=96, lmax=1280, nlmin=8, nlmax=40)
    meshes = [tbs.ragged_mesh(s) for s in shapes]
    lens = np.concatenate(shapes)
    res = {"workload": f"batch64_384000_seed{seed}", "nsys": len(meshes),
           "nelem": int(sum(m.nelem for m in meshes)), "nline": int(sum(m.nline for m in meshes)),
           "line_len_realised": [int(lens.min()), int(lens.max())],
           "shapes_note": "64 systems, lines per system U{8..40}, line lengths U[96,1280] drawn then scaled to "
                          "384,000 elements in all and clamped to [96,1280] (the independent draws total about "
                          "1.06M elements); realised lengths are line_len_realised"}
    bufs = []
    for i, m in enumerate(meshes):
        z = lambda k: ctx.array(size=k, zero=False)
        D, O1, O2, rhs = z(81 * m.nelem), z(81 * m.nintface), z(81 * m.nintface), z(9 * m.nelem)
        ctx.generate_device(1000 + i, m.nelem, m.nintface, args.shift, D, O1, O2, rhs)
        bufs.append((D, O1, O2, rhs))
    ctx.sync()
    refs = {}
    if oracle is not None:
        cores = host_cores()
        for i, m in enumerate(meshes):
            j = tbs.generate_host(1000 + i, 9, m.nelem, m.nintface, args.shift)
            refs[i] = (m, j, oracle.lij_run(m, j, 12, nthreads=cores))
    for prec, name in ((tbs.TBS_FP64, "fp64"), (tbs.TBS_FP32, "fp32_factor")):
        plans = [tbs.Plan(ctx, m, prec) for m in meshes]
        ctx.sync()
        t0 = time.perf_counter()
        tbs.factor_plans(ctx, plans, *[[b[k] for b in bufs] for k in range(3)])   # one launch for all
        t_factor = time.perf_counter() - t0
        b = tbs.Batch(ctx, plans)
        chosen, us = b.autotune(reps=3)
        tuning = b.tuning()
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
        st = b.states()
        if not (st[:, 1] == 1).all() or not (st[:, 0] == args.warmup + steps + 1).all():
            raise RuntimeError("a batched system stopped inside the timed sweeps")
        # time to 1e-10 from U = 0 (factorisation timed above, solve here)
        Us = [ctx.array(size=9 * m.nelem) for m in meshes]
        ctx.sync()
        t1 = time.perf_counter()
        hists, ns = b.solve([x[3] for x in bufs], Us, 100, rtol=1e-10, sweeps_per_check=2)
        ctx.sync()
        t_solve = time.perf_counter() - t1
        res[name] = {"sweeps_per_s": steps / t, "gbps_model": model / (t / steps) / 1e9,
                     "frac_hbm_peak": model / (t / steps) / 1e9 / HBM_PEAK_GBS, "bytes_per_sweep_model": model,
                     "autotune_variant": int(chosen), "autotune_tuple": tuning_tuple(tuning),
                     "autotune_pass_us_one_vs_two_streams": tuning["pass_us"], "factor_s": t_factor, "solve_to_1e-10_s": t_solve,
                     "time_to_1e-10_s": t_factor + t_solve, "max_sweeps_to_1e-10": int(max(ns)),
                     "worst_final_rel_residual": float(max(h[-1] / h[0] for h in hists))}
        if refs:
            if prec == tbs.TBS_FP64:
                # the same number of sweeps as the batch took: parity (1e-12)
                errs, hdev = [], []
                for i, (m, j, _) in refs.items():
                    r = oracle.lij_run(m, j, int(ns[i]), nthreads=host_cores())
                    errs.append(relerr(Us[i].numpy(), r["U"]))
                    hdev.append(hist_dev(hists[i], r["hist"]))
                res[name]["oracle_U_relerr_same_sweeps_worst"] = max(errs)
                res[name]["oracle_history_dev_worst"] = max(hdev)
            errs = [relerr(Us[i].numpy(), r["U"]) for i, (_, _, r) in refs.items()]
            res[name]["oracle_U_relerr_vs_converged_worst"] = max(errs)
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


def bench_serving(tbs, ctx, args, nreq=160, seed=7, oracle=None):
    """Continuous-batching server, FP32 factors / FP64 residual, tolerance
    1e-10.  Request sizes n log-uniform in [3840, 384000], line length L
    uniform in [64, 960], grid max(1, round(n/L)) x L.  Saturated throughput
    from a burst of nreq requests; then Poisson arrivals at 0.5, 0.8 and 0.95
    of it, latency = scheduled arrival -> result on the host.  With the
    oracle: every 20th request of the saturated run, its U against the FP64
    oracle's solution at the round-off floor."""
    import numpy as np
    rng = np.random.default_rng(seed)

    def draw(k):
        n = np.exp(rng.uniform(np.log(3840), np.log(384000), k))
        L = rng.integers(64, 961, k)
        NL = np.maximum(1, np.rint(n / L)).astype(int)
        return [(int(a), int(b)) for a, b in zip(NL, L)]

    def run(sizes, arrivals, keep=()):
        srv = tbs.Server(ctx, precision=tbs.TBS_FP32, rtol=1e-10, max_sweeps=100, sweeps_per_step=2)
        t0 = srv.info()["now"]
        rid_of, lat, nsw, status, phases, order = {}, [], [], [], [], []
        kept = {}
        k = 0
        while len(lat) < len(sizes):
            now = srv.info()["now"] - t0
            while k < len(sizes) and arrivals[k] <= now:
                NL, NK = sizes[k]
                rid_of[srv.submit(NL, NK, 5000 + k, args.shift)] = k
                k += 1
            srv.step()
            for rid in srv.poll():
                r = srv.result(rid, want_U=rid_of[rid] in keep)
                if rid_of[rid] in keep:
                    kept[rid_of[rid]] = r["U"]
                order.append(rid_of[rid])
                tl = srv.timeline(rid) - t0
                # arrival -> submitted -> admitted -> set-up done -> converged -> result
                phases.append([tl[0] - arrivals[rid_of[rid]], tl[1] - tl[0], tl[2] - tl[1], tl[3] - tl[2],
                               tl[5] - tl[3]])
                lat.append(r["t_done"] - t0 - arrivals[rid_of[rid]])
                nsw.append(r["nsweeps"])
                status.append(r["status"])
                srv.release(rid)
        wall = srv.info()["now"] - t0
        stats = srv.stats()
        srv.close()
        run.stats = stats
        run.kept = kept
        run.phases = np.array(phases)
        run.order = order
        return np.array(lat), wall, nsw, status

    sizes = draw(nreq)
    elems = sum(a * b for a, b in sizes)
    keep = set(range(0, nreq, 20)) if oracle is not None else set()
    lat, wall, nsw, st = run(sizes, np.zeros(nreq), keep)
    sat = nreq / wall
    out = {"workload": f"serve_{nreq}req_seed{seed}", "mean_elements": elems / nreq,
           "saturated_req_per_s": sat, "saturated_elements_per_s": elems / wall,
           "all_converged": bool(all(s == 1 for s in st)), "max_sweeps": int(max(nsw)),
           "saturated_host_phase_s": run.stats, "loads": {}}
    if keep:
        errs = {}
        for k, U in sorted(run.kept.items()):
            NL, NK = sizes[k]
            m = tbs.structured_mesh(NL, NK)
            j = tbs.generate_host(5000 + k, 9, m.nelem, m.nintface, args.shift)
            r = oracle.lij_run(m, j, 12, nthreads=host_cores())
            errs[f"{NL}x{NK}"] = relerr(U, r["U"])
        out["oracle_U_relerr_vs_converged"] = errs
        out["oracle_U_relerr_vs_converged_worst"] = max(errs.values())
    for frac in (0.5, 0.8, 0.95):
        sizes = draw(nreq)
        arr = np.cumsum(rng.exponential(1.0 / (frac * sat), nreq))
        lat, wall, nsw, st = run(sizes, arr)
        ph = run.phases
        tail = lat >= np.percentile(lat, 98)
        names = ("arrival_to_submit", "queued", "setup", "in_batch", "result_out")
        out["loads"][f"{frac:g}"] = {"offered_req_per_s": frac * sat, "achieved_req_per_s": nreq / wall,
                                     "p50_ms": float(np.percentile(lat, 50) * 1e3),
                                     "p99_ms": float(np.percentile(lat, 99) * 1e3),
                                     "all_converged": bool(all(s == 1 for s in st)),
                                     "host_phase_s": run.stats,
                                     "median_phase_ms": {k: float(np.median(ph[:, i]) * 1e3) for i, k in enumerate(names)},
                                     "tail2pct_phase_ms": {k: float(np.mean(ph[tail, i]) * 1e3) for i, k in enumerate(names)},
                                     "tail2pct_elements": [int(sizes[run.order[i]][0] * sizes[run.order[i]][1])
                                                           for i in np.argsort(lat)[-max(1, int(tail.sum())):]]}
    return out


def bench_dropin(tbs, ctx, m, D, O1, O2, rhs, sweeps=5):
    """The reference loop (src/main.cxx:180-229) on the drop-in kernels at the
    headline shape: copyAtoBjac + lineLU once, then per sweep solveDU, addAtoB,
    copyAtoB, lineRes, each kernel between its own pair of events.  GB/s per
    kernel on its algorithmic bytes; the sweep on the same byte model as the
    fused path (bytes_per_sweep_model)."""
    import numpy as np
    n, nif, nl = m.nelem, m.nintface, m.nline
    z = lambda k: ctx.array(size=k, zero=True)
    d = m.to_device(ctx)
    DLU, DinvC, Am = z(81 * n), z(81 * n), z(81 * n)
    U, dU, R = z(9 * n), z(9 * n), z(9 * n)
    ev = [ctx.event() for _ in range(10)]
    ctx.sync()
    ctx.record(ev[0])
    ctx.assembleTriBlocks(d, O1, O2, Am)
    ctx.record(ev[1])
    ctx.copyAtoBjac(n, D, DLU)
    ctx.lineLU(n, nif, d, DLU, O1, O2, DinvC)
    ctx.record(ev[2])
    ctx.copyAtoB(n, rhs, R)
    ctx.sync()
    t_asm, t_lu = ctx.elapsed_ms(ev[0], ev[1]), ctx.elapsed_ms(ev[1], ev[2])
    names = ("solveDU", "addAtoB", "copyAtoB", "lineRes")
    t = {k: [] for k in names}
    for sw in range(sweeps + 1):                    # the first sweep is a warm-up
        ctx.record(ev[0])
        ctx.solveDU(n, nif, d, DLU, DinvC, Am, dU, R)
        ctx.record(ev[1])
        ctx.addAtoB(n, dU, U)
        ctx.record(ev[2])
        ctx.copyAtoB(n, rhs, R)
        ctx.record(ev[3])
        ctx.lineRes(n, nif, d, D, O1, O2, U, R)
        ctx.record(ev[4])
        ctx.sync()
        if sw:
            for i, k in enumerate(names):
                t[k].append(ctx.elapsed_ms(ev[i], ev[i + 1]) * 1e-3)
    vec = 72.0 * n
    alg = {"solveDU": 648.0 * (3 * n - 2 * nl) + 2 * vec, "addAtoB": 3 * vec, "copyAtoB": 2 * vec,
           "lineRes": 648.0 * (n + 2 * nif) + 3 * vec}
    kern = {k: {"avg_ms": float(np.mean(t[k])) * 1e3, "alg_bytes": alg[k],
                "GBps": alg[k] / float(np.mean(t[k])) / 1e9} for k in names}
    t_sweep = sum(float(np.mean(t[k])) for k in names)
    model = tbs.model_bytes(m)
    out = {"workload": f"lij_{m.nline}x{n // m.nline}_dropin_reference_loop", "sweeps": sweeps,
           "sweeps_per_s": 1.0 / t_sweep, "gbps_model": model / t_sweep / 1e9,
           "frac_hbm_peak": model / t_sweep / 1e9 / HBM_PEAK_GBS,
           "gbps_reference_model": tbs.reference_model_bytes(m) / t_sweep / 1e9,
           "kernels": kern, "assembleTriBlocks_ms": t_asm, "copyAtoBjac_lineLU_ms": t_lu}
    for e in ev:
        ctx.destroy_event(e)
    for x in (DLU, DinvC, Am, U, dU, R, *d.values()):
        x.close()
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
    if args.cpu_baseline_only:
        cpu_baseline_child(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is designed "
                                   "(DESIGN.md) but not built", "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import triblocksolve_amd as tbs

    ctx = tbs.Context(0, 9)
    if args.section:
        fn = bench_batched if args.section == "batched" else bench_serving
        print(json.dumps({args.section: fn(tbs, ctx, args)}))
        return
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
        tuned = plan.autotune(reps=3) + (plan.tuning(),)
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
        # hipGraph; the first launch captures and is untimed.  The iterate is
        # restarted before the timed replay, and every sweep of it must have
        # run (a system that stops early would make its kernels return at once)
        ctx.sync()
        plan.launch_sweeps(steps, use_graph=True)
        plan.begin(rhs, None)
        ctx.sync()
        t1 = time.perf_counter()
        plan.launch_sweeps(steps, use_graph=True)
        ctx.sync()
        t_graph = time.perf_counter() - t1
        h = plan.history()
        if len(h) != steps + 1 or not np.all(np.isfinite(h)):
            raise RuntimeError(f"timed replay recorded {len(h)} of {steps + 1} residual norms")
        plan.close()
        return dict(wall=wall, thomas=th, residual=rs, norm=nm, info=info, t_factor=t_factor, t_graph=t_graph,
                    tuned=tuned)

    def oracle_parity(ref, iters=8):
        """FP64 plan, `iters` sweeps from zero, against the oracle's same sweeps."""
        plan = tbs.Plan(ctx, m, tbs.TBS_FP64)
        plan.factor(D, O1, O2)
        U = ctx.array(size=9 * n)
        hist, ns = plan.solve(rhs, U, iters, 0.0)
        plan.close()
        r = {"sweeps": int(ns), "U_relerr": relerr(U.numpy(), ref["U"]), "history_dev": hist_dev(hist, ref["hist"])}
        U.close()
        return r

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

    log("fp64 timed sweeps")
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
                   "GBps": info["thomas_bytes"] / t_th / 1e9,
                   # the forward intermediate d' written and read back (not in
                   # the algorithmic bytes): the rate of all bytes moved
                   "dprime_roundtrip_bytes": 2 * 8 * 9 * n,
                   "GBps_incl_dprime": (info["thomas_bytes"] + 2 * 8 * 9 * n) / t_th / 1e9},
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
        "autotune": {"chosen_variant": r64["tuned"][0], "thomas_us_by_variant": r64["tuned"][1],
                     "tuple": tuning_tuple(r64["tuned"][2]),
                     "residual_us_by_waves_per_wg": residual_by_rw(r64["tuned"][2])},
    }
    log("stream ceiling")
    ceil = ctx.stream_ceiling()
    out["achievable_gbps"] = ceil
    roof["frac_of_copy_ceiling"] = kern[dom]["GBps"] / ceil["copy_gbps"]
    out["frac_of_copy_ceiling"] = gbs / ceil["copy_gbps"]
    oracle, ref = None, None
    if not args.no_oracle:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_py import Oracle
        oracle = Oracle()
        log("oracle on the headline system")
        jh = tbs.generate_host(args.seed, 9, n, nif, args.shift)
        t0 = time.perf_counter()
        ref = oracle.lij_run(m, jh, 8, nthreads=host_cores())
        del jh
        par = oracle_parity(ref, 8)
        out["oracle"] = {"what": "FP64 oracle (the reference's CPU path restated, tests/oracle_py.py) on the same "
                                 "seeded system, 8 sweeps from zero, outside every timed region",
                         "oracle_s": time.perf_counter() - t0, "fp64_same_sweeps": par,
                         "tolerance": "history |h_k/h_0 - h'_k/h'_0| <= 1e-12 max(h'_k/h'_0, 1e-4); U <= 1e-12"}
    def section(name, fn):
        """An extra section: its failure is recorded in the JSON line (with the
        traceback on stderr) instead of ending the run before the line."""
        log(name)
        try:
            return fn()
        except Exception as e:                      # noqa: BLE001
            traceback.print_exc(file=sys.stderr)
            return {"error": f"{type(e).__name__}: {e}"}

    if not args.no_extras:
        def fp32_section():
            r32 = timed_sweeps(tbs.TBS_FP32, args.steps, args.warmup)
            i32 = r32["info"]
            return {
                "sweeps_per_s": args.steps / r32["t_graph"],
                "gbps_model": i32["alg_bytes_per_sweep"] / (r32["t_graph"] / args.steps) / 1e9,
                "thomas_ms": float(r32["thomas"].mean()) * 1e3,
                "residual_ms": float(r32["residual"].mean()) * 1e3,
                "eager_sweeps_per_s": args.steps / r32["wall"],
            }

        def tol_section():
            a = time_to_tol(tbs.TBS_FP64)
            b = time_to_tol(tbs.TBS_FP32)
            ua, ub = a["U"].numpy(), b["U"].numpy()
            err = float(np.linalg.norm(ub - ua) / np.linalg.norm(ua))
            for d_, u_ in ((a, ua), (b, ub)):
                d_.pop("U").close()
                if ref is not None:
                    # against the oracle's solution at the round-off floor (8 sweeps)
                    d_["oracle_U_relerr_vs_converged"] = relerr(u_, ref["U"])
            return {"fp64": a, "fp32_factor": b, "fp32_vs_fp64_U_relerr": err}

        out["fp32_factor"] = section("fp32-factor timed sweeps", fp32_section)
        out["time_to_1e-10"] = section("time to 1e-10", tol_section)
        out["dropin"] = section("drop-in reference loop", lambda: bench_dropin(tbs, ctx, m, D, O1, O2, rhs))
        out["gmres"] = section("gmres", lambda: bench_gmres(tbs, ctx, m, args))
        out["batched"] = section("batched", lambda: bench_batched(tbs, ctx, args, oracle=oracle))
        out["serving"] = section("serving", lambda: bench_serving(tbs, ctx, args, oracle=oracle))
        out["block_sizes"] = section("block sizes", lambda: bench_block_sizes(tbs, args))
    if not args.no_cpu:
        cb = section("cpu baseline (child process)", lambda: cpu_baseline(args))
        if "error" in cb:
            out["cpu_baseline"] = cb
        else:
            out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
            out["cpu_baseline_detail"] = {k: v for k, v in cb.items() if k not in out["cpu_baseline"]}
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
