// main.cxx -- triblock.exe <compute_mode> <device_id> <block_size> [options]
//
// The reference driver (src/main.cxx:13-269) on the HIP library: read
// gpuline.mesh.data.bin / gpuline.jacobian.data.bin from the working
// directory (or generate the seeded structured workload with --synthetic),
// factor (copyAtoBjac + lineLU), then `iters` LIJ sweeps of solveDU, U += dU,
// R = rhs, lineRes, with per-phase device timing and the reference's GB/s
// byte model.  Additions: the residual norm after every sweep, and --fused to
// run the same iteration through the graph-captured fused path.
//
// Options: --synthetic NLxNK  --shift S  --seed N  --iters N  --fused
//          --precision fp64|fp32  --rtol R  --write-data  --write-resized FILE
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "Jacobian.hxx"
#include "Mesh.hxx"
#include "Platform.hxx"

namespace {

struct Timer {
    hipEvent_t a, b;
    hipStream_t s;
    explicit Timer(void *stream) : s((hipStream_t)stream) {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
    }
    ~Timer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    void start() { (void)hipEventRecord(a, s); }
    double stop() {
        (void)hipEventRecord(b, s);
        (void)hipEventSynchronize(b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        return ms * 1e-3;
    }
};

void usage() {
    printf("Arguments:\n"
           "  compute_mode: 1=HIP (the only mode built; 0=Serial, 2=CUDA, 3=OpenCL, 4=OpenMP,\n"
           "                5=DPC++, 6=Metal of the reference are not part of this build)\n"
           "  device_id:    Device ID on node\n"
           "  block_size:   Size of the matrix sub-block (e.g. 9 for 9x9 blocks)\n"
           "Options: --synthetic NLxNK --shift S --seed N --iters N --fused --precision fp64|fp32\n"
           "         --rtol R --write-data --write-resized FILE\n");
}

}  // namespace

int main(int argc, char **argv) {
    int compute_mode = HIP_MODE, device_id = 0, block_size = 9, iters = 30;
    int NL = 0, NK = 0, seed = 1, precision = TBS_FP64;
    double shift = 20.0, rtol = 0.0;
    bool fused = false, write_data = false;
    std::string resized_path;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); exit(2); }
            return argv[++i];
        };
        if (a == "--help" || a == "-help") { usage(); return 0; }
        else if (a == "--synthetic") { std::string v = next(); if (sscanf(v.c_str(), "%dx%d", &NL, &NK) != 2) { usage(); return 2; } }
        else if (a == "--shift") shift = std::stod(next());
        else if (a == "--seed") seed = std::stoi(next());
        else if (a == "--iters") iters = std::stoi(next());
        else if (a == "--rtol") rtol = std::stod(next());
        else if (a == "--fused") fused = true;
        else if (a == "--write-data") write_data = true;
        else if (a == "--write-resized") resized_path = next();
        else if (a == "--precision") precision = (next() == "fp32") ? TBS_FP32 : TBS_FP64;
        else pos.push_back(a);
    }
    if (pos.size() > 0) compute_mode = std::stoi(pos[0]);
    if (pos.size() > 1) device_id = std::stoi(pos[1]);
    if (pos.size() > 2) block_size = std::stoi(pos[2]);
    printf(" TRIDIAGONAL BLOCK SOLVER: Block size: %d, Compute mode: %d, Device ID: %d\n", block_size,
           compute_mode, device_id);
    try {
        // the inputs are read and checked before the device is touched; the
        // platform outlives the mesh and Jacobian (their device arrays are
        // freed through it)
        std::unique_ptr<Platform> gpu_holder;
        Mesh mesh;
        Jacobian Jac;
        if (NL > 0) {
            if (!mesh.fromStructured(NL, NK, block_size)) { fprintf(stderr, "bad grid\n"); return 1; }
            if (!Jac.fromGenerator(mesh, (uint64_t)seed, shift)) { fprintf(stderr, "generator failed\n"); return 1; }
            if (write_data) {
                mesh.toFile("gpuline.mesh.data.bin");
                Jac.toFile("gpuline.jacobian.data.bin");
            }
        } else {
            if (!mesh.fromFile()) return 1;
            if (!Jac.fromFile(mesh.nlineelem)) return 1;
            if (Jac.nvar != mesh.nvar || Jac.nelem != mesh.nelem || Jac.nintface != mesh.nintface) {
                fprintf(stderr,
                        "ERROR: the Jacobian (nvar %d, nelem %d, nintface %d) is not the mesh's system "
                        "(nvar %d, nelem %d, nintface %d)\n",
                        Jac.nvar, Jac.nelem, Jac.nintface, mesh.nvar, mesh.nelem, mesh.nintface);
                return 1;
            }
        }
        gpu_holder.reset(new Platform(compute_mode, device_id, block_size));
        Platform &gpu = *gpu_holder;
        Jac.assembleTriBlocks(mesh);
        Jac.resizeBlockSize(block_size);
        if (Jac.nvar != mesh.nvar) { mesh.nvar = Jac.nvar; Jac.assembleTriBlocks(mesh); }
        if (!resized_path.empty() && !Jac.toFile(resized_path)) { fprintf(stderr, "cannot write %s\n", resized_path.c_str()); return 1; }
        mesh.setupDevice(gpu);
        mesh.toDevice();
        Jac.setupDevice(gpu);
        Jac.toDevice();
        gpu.finish();
        printf("GPU Memory Allocated (MB): %.1f\n", gpu.memoryAllocated() / 1024. / 1024.);
        const int nv = Jac.nvar, n = mesh.nelem;
        const long nvec = (long)nv * n;
        Timer t(gpu.stream());
        std::vector<double> hist;
        double LU_time = 0, dU_time = 0, cp_time = 0, LR_time = 0;
        if (!fused) {
            t.start();
            tbs_check(tbs_copyAtoBjac(gpu.ctx, n, Jac.o_jacD, Jac.o_jacDLU), "copyAtoBjac");
            tbs_check(tbs_lineLU(gpu.ctx, n, mesh.nintface, mesh.nline, mesh.o_fc, mesh.o_linesize,
                                 mesh.o_linepoint, mesh.o_lines, mesh.o_lineface, Jac.o_jacDLU,
                                 Jac.o_jacO1, Jac.o_jacO2, Jac.o_jacDinvC), "lineLU");
            LU_time = t.stop();
            printf("   LU Time: %f\n", LU_time);
            t.start();
            tbs_check(tbs_copyAtoB(gpu.ctx, n, Jac.o_rhs, Jac.o_res), "copyAtoB");
            cp_time += t.stop();
            double r0;
            tbs_check(tbs_norm2(gpu.ctx, nvec, Jac.o_res, &r0), "norm");
            hist.push_back(r0);
            for (int p = 0; p < iters; ++p) {
                t.start();
                tbs_check(tbs_triblock_solveDU(gpu.ctx, n, mesh.nintface, mesh.eftot, mesh.nline,
                                               mesh.nlineelem, mesh.o_linesize, mesh.o_linepoint,
                                               mesh.o_lines, Jac.o_jacDLU, Jac.o_jacDinvC, Jac.o_A,
                                               Jac.o_dU, Jac.o_res), "solveDU");
                dU_time += t.stop();
                t.start();
                tbs_check(tbs_addAtoB(gpu.ctx, n, Jac.o_dU, Jac.o_U), "addAtoB");
                tbs_check(tbs_copyAtoB(gpu.ctx, n, Jac.o_rhs, Jac.o_res), "copyAtoB");
                cp_time += t.stop();
                t.start();
                tbs_check(tbs_triblock_lineRes(gpu.ctx, n, mesh.nintface, mesh.eftot, mesh.nline,
                                               mesh.nlineelem, mesh.o_epoint, mesh.o_ef, mesh.o_fc,
                                               mesh.o_linesize, mesh.o_linepoint, mesh.o_lines,
                                               Jac.o_jacD, Jac.o_jacO1, Jac.o_jacO2, Jac.o_U,
                                               Jac.o_res), "lineRes");
                LR_time += t.stop();
                double r;
                tbs_check(tbs_norm2(gpu.ctx, nvec, Jac.o_res, &r), "norm");
                hist.push_back(r);
                if (rtol > 0 && r <= rtol * hist[0]) { iters = p + 1; break; }
            }
            // the reference's byte model (src/main.cxx:149-175)
            const double bs = (double)nv * nv;
            const double duMem = (8.0 * (3 * bs * n + 4.0 * nv * n + nv * n) +
                                  4.0 * (2.0 * mesh.nline + 1 + 2.0 * mesh.nlineelem)) * iters / 1e9;
            const double cpMem = 5.0 * nv * n * 8.0 * (iters + 0.4) / 1e9;
            const double lrMem = (8.0 * (bs * n + 2 * bs * mesh.nintface + 7.0 * nv * n) +
                                  4.0 * (n + 1.0 + mesh.eftot + 2.0 * mesh.nintface + 2.0 * mesh.nline + 1 +
                                         2.0 * mesh.nlineelem)) * iters / 1e9;
            const double tot = dU_time + cp_time + LR_time;
            printf("-----------------------------------------\n");
            printf("[hip] Time: %f %f GB/s\n", tot, (duMem + cpMem + lrMem) / tot);
            printf("   du Time: %f %f GB/s\n", dU_time, duMem / dU_time);
            printf("U+=du Time: %f %f GB/s\n", cp_time, cpMem / cp_time);
            printf("linRes Time: %f %f GB/s\n", LR_time, lrMem / LR_time);
        } else {
            tbs_mesh_host mh = mesh.host();
            tbs_plan *plan = nullptr;
            tbs_check(tbs_plan_create(gpu.ctx, &mh, precision, &plan), "plan_create");
            t.start();
            tbs_check(tbs_plan_factor(plan, Jac.o_jacD, Jac.o_jacO1, Jac.o_jacO2), "plan_factor");
            LU_time = t.stop();
            printf("   LU Time: %f\n", LU_time);
            hist.assign(iters + 1, 0.0);
            int ns = 0;
            auto c0 = std::chrono::steady_clock::now();
            tbs_check(tbs_plan_solve(plan, Jac.o_rhs, Jac.o_U, iters, rtol, hist.data(), &ns, Jac.o_res),
                      "plan_solve");
            const double ts = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
            hist.resize(ns + 1);
            iters = ns;
            tbs_plan_info info;
            tbs_plan_info_get(plan, &info);
            printf("-----------------------------------------\n");
            printf("[fused %s] Time: %f  %.1f sweeps/s  %.1f GB/s (model %.3f GB/sweep)\n",
                   precision == TBS_FP32 ? "fp32-factor" : "fp64", ts, ns / ts,
                   info.alg_bytes_per_sweep * ns / ts / 1e9, info.alg_bytes_per_sweep / 1e9);
            tbs_plan_destroy(plan);
        }
        printf("-----------------------------------------\n");
        printf("Iterations: %d\nTotal Time: %f\n", iters, dU_time + cp_time + LR_time + LU_time);
        for (size_t k = 0; k < hist.size(); ++k)
            printf("  ||R_%zu|| = %.17e  (%.3e relative)\n", k, hist[k], hist[k] / hist[0]);
        Jac.fromDevice();
    } catch (const std::exception &e) {
        fprintf(stderr, "ERROR: %s\n", e.what());
        return 1;
    }
    return 0;
}
