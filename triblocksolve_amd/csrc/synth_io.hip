// synth_io.hip -- seeded synthetic workloads and the reference's binary
// formats.
//
// Formats (both Fortran stream-unformatted, little-endian):
//   gpuline.mesh.data.bin: int32 header nvar,nelem,nintface,eftot,nline,
//     nlineelem, then epoint[nelem+1], ef[eftot], fc[2*nintface],
//     linesize[nline], linepoint[nline+1], lines[nlineelem],
//     lineface[nlineelem]; every index 1-based in the file, shifted to 0-based
//     on read (src/F90/read_mesh_data.F90:27-64, src/Mesh.cxx:11-73).
//   gpuline.jacobian.data.bin: int32 nvar,nelem,nintface, then FP64
//     jacD[nvar*nvar*nelem], jacO1/jacO2[nvar*nvar*nintface], rhs and U0
//     [nvar*nelem] (src/F90/read_jacobian_data.F90:25-31).
// Unlike the reference readers, these check that the file size matches the
// header before reading anything.
//
// Generator: counter-based, so the host and device variants give identical
// bits for any element range.  Array ids: 1 = jacD, 2 = jacO1, 3 = jacO2,
// 4 = rhs.  value(seed, id, i) = a + (b - a) * u with
//   key = mix64(seed*G + id*H + 1),  u = (mix64(key + (i+1)*G) >> 11) * 2^-53
// each product and sum rounded separately (no contraction); jacD's diagonal
// then gets + shift.  Structured mesh: element e = l*NK + k (line l, position
// k), interior faces numbered in element order (in-line face (e+1,e) first,
// then cross face (e+NK,e)), ef = [S,E,N,W] per element, boundary faces
// numbered S (k order), E (l order), N (k order), W (l order) with file ids
// -1, -2, ...; this reproduces data/ex05/gpuline.mesh.data.bin exactly.
#include "tbs_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sys/stat.h>

namespace {

constexpr uint64_t kG = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kH = 0xD1B54A32D192ED03ull;

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t stream_key(uint64_t seed, uint64_t id) {
    return mix64(seed * kG + id * kH + 1ull);
}

__host__ __device__ inline double uniform_at(uint64_t key, uint64_t i, double a, double b) {
#pragma clang fp contract(off)
    const double u = (double)(mix64(key + (i + 1ull) * kG) >> 11) * 0x1.0p-53;
    const double w = b - a;
    const double t = w * u;
    return a + t;
}

__host__ __device__ inline double jacD_at(uint64_t key, uint64_t i, int nv, double shift) {
#pragma clang fp contract(off)
    double v = uniform_at(key, i, -1.0, 1.0);
    const int q = (int)(i % (uint64_t)(nv * nv));
    if (q % (nv + 1) == 0) v = v + shift;
    return v;
}

__global__ void gen_kernel(uint64_t seed, int nv, long nD, long nO, long nR, double shift,
                           double *jacD, double *O1, double *O2, double *rhs) {
    const uint64_t kD = stream_key(seed, 1), k1 = stream_key(seed, 2);
    const uint64_t k2 = stream_key(seed, 3), kR = stream_key(seed, 4);
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nD; i += stride)
        jacD[i] = jacD_at(kD, (uint64_t)i, nv, shift);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nO; i += stride) {
        O1[i] = uniform_at(k1, (uint64_t)i, -0.1, 0.1);
        O2[i] = uniform_at(k2, (uint64_t)i, -0.1, 0.1);
    }
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nR; i += stride)
        rhs[i] = uniform_at(kR, (uint64_t)i, -1.0, 1.0);
}

int file_size(const char *path, long *sz) {
    struct stat st;
    if (stat(path, &st) != 0) {
        tbs::set_error("could not find %s", path);
        return TBS_ERR_IO;
    }
    *sz = (long)st.st_size;
    return TBS_OK;
}

long mesh_file_bytes(const int h[6]) {
    return 4L * (6L + (h[1] + 1L) + h[3] + 2L * h[2] + h[4] + (h[4] + 1L) + 2L * h[5]);
}

long jac_file_bytes(const int h[3]) {
    const long bs = (long)h[0] * h[0];
    return 12L + 8L * (bs * h[1] + 2L * bs * h[2] + 2L * (long)h[0] * h[1]);
}

bool read_n(FILE *fp, void *dst, size_t bytes) { return fread(dst, 1, bytes, fp) == bytes; }

}  // namespace

extern "C" {

int tbs_structured_sizes(int NL, int NK, int *nelem, int *nintface, int *eftot,
                         int *nline, int *nlineelem) {
    TBS_REQUIRE(NL >= 1 && NK >= 1, TBS_ERR_ARG, "grid %d x %d must be >= 1 x 1", NL, NK);
    TBS_REQUIRE((long)NL * NK * 4 < (1L << 31), TBS_ERR_ARG, "grid %d x %d overflows int32", NL, NK);
    const int n = NL * NK;
    if (nelem) *nelem = n;
    if (nintface) *nintface = NL * (NK - 1) + (NL - 1) * NK;
    if (eftot) *eftot = 4 * n;
    if (nline) *nline = NL;
    if (nlineelem) *nlineelem = n;
    return TBS_OK;
}

int tbs_structured_mesh(int NL, int NK, int *epoint, int *ef, int *fc, int *linesize,
                        int *linepoint, int *lines, int *lineface) {
    int n, nif;
    int rc = tbs_structured_sizes(NL, NK, &n, &nif, nullptr, nullptr, nullptr);
    if (rc) return rc;
    TBS_REQUIRE(epoint && ef && (fc || nif == 0) && linesize && linepoint && lines && lineface, TBS_ERR_ARG,
                "null mesh array");
    // face ids of the in-line face (e+1,e) and cross face (e+NK,e) created at e
    std::vector<int> f_in(n, -1), f_cross(n, -1);
    int f = 0;
    for (int e = 0; e < n; ++e) {
        const int l = e / NK, k = e % NK;
        if (k < NK - 1) { f_in[e] = f; fc[2 * f] = e + 1; fc[2 * f + 1] = e; ++f; }
        if (l < NL - 1) { f_cross[e] = f; fc[2 * f] = e + NK; fc[2 * f + 1] = e; ++f; }
    }
    // boundary ids, 0-based after the reader's shift: file id -(b+1) -> -(b+2)
    auto bnd = [](int b) { return -(b + 1) - 1; };
    const int bS = 0, bE = NK, bN = NK + NL, bW = NK + NL + NK;
    for (int e = 0; e < n; ++e) {
        const int l = e / NK, k = e % NK;
        epoint[e] = 4 * e;
        ef[4 * e + 0] = (l > 0) ? f_cross[e - NK] : bnd(bS + k);
        ef[4 * e + 1] = (k < NK - 1) ? f_in[e] : bnd(bE + l);
        ef[4 * e + 2] = (l < NL - 1) ? f_cross[e] : bnd(bN + k);
        ef[4 * e + 3] = (k > 0) ? f_in[e - 1] : bnd(bW + l);
        lines[e] = e;
        lineface[e] = ef[4 * e + 3];
    }
    epoint[n] = 4 * n;
    for (int l = 0; l < NL; ++l) { linesize[l] = NK; linepoint[l] = l * NK; }
    linepoint[NL] = n;
    return TBS_OK;
}

int tbs_ragged_sizes(int nline, const int *lens, int *nelem, int *nintface, int *eftot) {
    TBS_REQUIRE(nline >= 1 && lens, TBS_ERR_ARG, "bad ragged mesh arguments");
    long n = 0, nif = 0;
    for (int l = 0; l < nline; ++l) {
        TBS_REQUIRE(lens[l] >= 1, TBS_ERR_ARG, "line %d has length %d", l, lens[l]);
        n += lens[l];
        nif += lens[l] - 1;
        if (l + 1 < nline) nif += std::min(lens[l], lens[l + 1]);
    }
    TBS_REQUIRE(4 * n < (1L << 31), TBS_ERR_ARG, "ragged mesh overflows int32");
    if (nelem) *nelem = (int)n;
    if (nintface) *nintface = (int)nif;
    if (eftot) *eftot = (int)(4 * n);
    return TBS_OK;
}

// Lines of uneven length stacked like the structured grid: element (l, k) =
// start[l] + k, in-line faces (e+1, e), cross faces (start[l+1]+k, e) when
// line l+1 has a position k; ef = [S, E, N, W]; boundary faces numbered by
// side (S, E, N, W), each in element order.  Equal lengths give exactly the
// structured mesh.
int tbs_ragged_mesh(int nline, const int *lens, int *epoint, int *ef, int *fc, int *linesize,
                    int *linepoint, int *lines, int *lineface) {
    int n, nif, eft;
    int rc = tbs_ragged_sizes(nline, lens, &n, &nif, &eft);
    if (rc) return rc;
    TBS_REQUIRE(epoint && ef && (fc || nif == 0) && linesize && linepoint && lines && lineface, TBS_ERR_ARG,
                "null mesh array");
    std::vector<int> start(nline + 1, 0);
    for (int l = 0; l < nline; ++l) start[l + 1] = start[l] + lens[l];
    std::vector<int> f_in(n, -1), f_cross(n, -1);
    int f = 0;
    for (int l = 0; l < nline; ++l)
        for (int k = 0; k < lens[l]; ++k) {
            const int e = start[l] + k;
            if (k < lens[l] - 1) { f_in[e] = f; fc[2 * f] = e + 1; fc[2 * f + 1] = e; ++f; }
            if (l + 1 < nline && k < lens[l + 1]) {
                f_cross[e] = f; fc[2 * f] = start[l + 1] + k; fc[2 * f + 1] = e; ++f;
            }
        }
    // boundary ids by side, element order within a side (0-based: -(b+2))
    int b = 0;
    std::vector<int> bS(n, 0), bE(n, 0), bN(n, 0), bW(n, 0);
    for (int pass = 0; pass < 4; ++pass)
        for (int l = 0; l < nline; ++l)
            for (int k = 0; k < lens[l]; ++k) {
                const int e = start[l] + k;
                if (pass == 0 && !(l > 0 && k < lens[l - 1])) bS[e] = -(b++) - 2;
                if (pass == 1 && k == lens[l] - 1) bE[e] = -(b++) - 2;
                if (pass == 2 && f_cross[e] < 0) bN[e] = -(b++) - 2;
                if (pass == 3 && k == 0) bW[e] = -(b++) - 2;
            }
    for (int l = 0; l < nline; ++l) {
        linesize[l] = lens[l];
        linepoint[l] = start[l];
        for (int k = 0; k < lens[l]; ++k) {
            const int e = start[l] + k;
            epoint[e] = 4 * e;
            ef[4 * e + 0] = (l > 0 && k < lens[l - 1]) ? f_cross[start[l - 1] + k] : bS[e];
            ef[4 * e + 1] = (k < lens[l] - 1) ? f_in[e] : bE[e];
            ef[4 * e + 2] = (f_cross[e] >= 0) ? f_cross[e] : bN[e];
            ef[4 * e + 3] = (k > 0) ? f_in[e - 1] : bW[e];
            lines[e] = e;
            lineface[e] = ef[4 * e + 3];
        }
    }
    epoint[n] = 4 * n;
    linepoint[nline] = n;
    return TBS_OK;
}

// Seeded batch shapes: system s gets nl_s ~ U{nlmin..nlmax} lines, line
// lengths ~ U{lmin..lmax}; when total > 0 every length is scaled by
// total / sum, rounded, clamped to [lmin, lmax], and the remainder is
// settled one element at a time from the last line backwards.
int tbs_batch_shapes(uint64_t seed, int nsys, long total, int lmin, int lmax, int nlmin, int nlmax,
                     int *nlines, int *lens, int cap, int *nlens) {
    TBS_REQUIRE(nsys >= 1 && lmin >= 1 && lmax >= lmin && nlmin >= 1 && nlmax >= nlmin && nlines &&
                    nlens, TBS_ERR_ARG, "bad batch shape arguments");
    uint64_t st = seed * kG + 0x5851F42D4C957F2Dull;
    auto next = [&]() { st += kG; return mix64(st); };
    std::vector<long> len;
    for (int s = 0; s < nsys; ++s) {
        nlines[s] = nlmin + (int)(next() % (uint64_t)(nlmax - nlmin + 1));
        for (int l = 0; l < nlines[s]; ++l) len.push_back(lmin + (long)(next() % (uint64_t)(lmax - lmin + 1)));
    }
    const long nl = (long)len.size();
    TBS_REQUIRE(nl <= cap || !lens, TBS_ERR_ARG, "lens capacity %d < %ld lines", cap, nl);
    if (total > 0) {
        TBS_REQUIRE(total >= nl * (long)lmin && total <= nl * (long)lmax, TBS_ERR_ARG,
                    "total %ld elements not reachable with %ld lines of [%d, %d]", total, nl, lmin, lmax);
        double sum = 0;
        for (long v : len) sum += (double)v;
        const double f = (double)total / sum;
        long acc = 0;
        for (long &v : len) {
            v = std::min<long>(lmax, std::max<long>(lmin, std::lround((double)v * f)));
            acc += v;
        }
        long diff = total - acc;
        while (diff != 0) {
            bool moved = false;
            for (long i = nl - 1; i >= 0 && diff != 0; --i) {
                if (diff > 0 && len[i] < lmax) { ++len[i]; --diff; moved = true; }
                else if (diff < 0 && len[i] > lmin) { --len[i]; ++diff; moved = true; }
            }
            TBS_REQUIRE(moved, TBS_ERR_ARG, "cannot settle batch total");
        }
    }
    *nlens = (int)nl;
    if (lens) for (long i = 0; i < nl; ++i) lens[i] = (int)len[i];
    return TBS_OK;
}

int tbs_generate_host(uint64_t seed, int nvar, int nelem, int nintface, double shift,
                      double *jacD, double *jacO1, double *jacO2, double *rhs) {
    TBS_REQUIRE(nvar >= 1 && nvar <= 64 && nelem >= 0 && nintface >= 0, TBS_ERR_ARG,
                "bad generator sizes");
    const uint64_t kD = stream_key(seed, 1), k1 = stream_key(seed, 2);
    const uint64_t k2 = stream_key(seed, 3), kR = stream_key(seed, 4);
    const long bs = (long)nvar * nvar;
    if (jacD) for (long i = 0; i < bs * nelem; ++i) jacD[i] = jacD_at(kD, (uint64_t)i, nvar, shift);
    if (jacO1) for (long i = 0; i < bs * nintface; ++i) jacO1[i] = uniform_at(k1, (uint64_t)i, -0.1, 0.1);
    if (jacO2) for (long i = 0; i < bs * nintface; ++i) jacO2[i] = uniform_at(k2, (uint64_t)i, -0.1, 0.1);
    if (rhs) for (long i = 0; i < (long)nvar * nelem; ++i) rhs[i] = uniform_at(kR, (uint64_t)i, -1.0, 1.0);
    return TBS_OK;
}

int tbs_generate_device(tbs_ctx *ctx, uint64_t seed, int nvar, int nelem, int nintface,
                        double shift, double *jacD, double *jacO1, double *jacO2, double *rhs) {
    int rc = tbs::use_ctx(ctx);
    if (rc) return rc;
    TBS_REQUIRE(nvar >= 1 && nvar <= 64 && nelem >= 0 && nintface >= 0, TBS_ERR_ARG,
                "bad generator sizes");
    TBS_REQUIRE(jacD && jacO1 && jacO2 && rhs, TBS_ERR_ARG, "null output array");
    const long bs = (long)nvar * nvar;
    hipLaunchKernelGGL(gen_kernel, dim3(8192), dim3(256), 0, ctx->stream, seed, nvar,
                       bs * nelem, bs * nintface, (long)nvar * nelem, shift, jacD, jacO1,
                       jacO2, rhs);
    TBS_LAUNCHED("gen_kernel");
    return TBS_OK;
}

int tbs_read_mesh_header(const char *path, int sizes[6]) {
    TBS_REQUIRE(path && sizes, TBS_ERR_ARG, "null argument");
    long sz;
    int rc = file_size(path, &sz);
    if (rc) return rc;
    FILE *fp = fopen(path, "rb");
    TBS_REQUIRE(fp, TBS_ERR_IO, "could not open %s", path);
    bool ok = read_n(fp, sizes, 6 * sizeof(int));
    fclose(fp);
    TBS_REQUIRE(ok, TBS_ERR_IO, "%s: truncated header", path);
    for (int i = 0; i < 6; ++i)
        TBS_REQUIRE(sizes[i] >= 0, TBS_ERR_IO, "%s: negative size in header", path);
    TBS_REQUIRE(mesh_file_bytes(sizes) == sz, TBS_ERR_IO,
                "%s: header implies %ld bytes, file has %ld", path, mesh_file_bytes(sizes), sz);
    return TBS_OK;
}

int tbs_read_mesh(const char *path, int *epoint, int *ef, int *fc, int *linesize,
                  int *linepoint, int *lines, int *lineface) {
    int h[6];
    int rc = tbs_read_mesh_header(path, h);
    if (rc) return rc;
    FILE *fp = fopen(path, "rb");
    TBS_REQUIRE(fp, TBS_ERR_IO, "could not open %s", path);
    int hh[6];
    bool ok = read_n(fp, hh, sizeof(hh));
    // shift: every index array, not the linesize counts (read_mesh_data.F90:40-64)
    struct { int *p; long n; int shift; } arr[7] = {{epoint, h[1] + 1L, 1}, {ef, h[3], 1},
                                                   {fc, 2L * h[2], 1}, {linesize, h[4], 0},
                                                   {linepoint, h[4] + 1L, 1}, {lines, h[5], 1},
                                                   {lineface, h[5], 1}};
    for (auto &a : arr) {
        if (!ok) break;
        ok = read_n(fp, a.p, a.n * sizeof(int));
        for (long i = 0; ok && i < a.n; ++i) a.p[i] -= a.shift;
    }
    fclose(fp);
    TBS_REQUIRE(ok, TBS_ERR_IO, "%s: short read", path);
    return TBS_OK;
}

int tbs_write_mesh(const char *path, const int sizes[6], const int *epoint, const int *ef,
                   const int *fc, const int *linesize, const int *linepoint, const int *lines,
                   const int *lineface) {
    TBS_REQUIRE(path && sizes, TBS_ERR_ARG, "null argument");
    FILE *fp = fopen(path, "wb");
    TBS_REQUIRE(fp, TBS_ERR_IO, "could not create %s", path);
    bool ok = fwrite(sizes, sizeof(int), 6, fp) == 6;
    struct { const int *p; long n; int shift; } arr[7] = {
        {epoint, sizes[1] + 1L, 1}, {ef, sizes[3], 1}, {fc, 2L * sizes[2], 1}, {linesize, sizes[4], 0},
        {linepoint, sizes[4] + 1L, 1}, {lines, sizes[5], 1}, {lineface, sizes[5], 1}};
    std::vector<int> tmp;
    for (auto &a : arr) {
        if (!ok) break;
        tmp.assign(a.p, a.p + a.n);
        for (auto &v : tmp) v += a.shift;
        ok = fwrite(tmp.data(), sizeof(int), a.n, fp) == (size_t)a.n;
    }
    ok = (fclose(fp) == 0) && ok;
    TBS_REQUIRE(ok, TBS_ERR_IO, "%s: write failed", path);
    return TBS_OK;
}

int tbs_read_jacobian_header(const char *path, int sizes[3]) {
    TBS_REQUIRE(path && sizes, TBS_ERR_ARG, "null argument");
    long sz;
    int rc = file_size(path, &sz);
    if (rc) return rc;
    FILE *fp = fopen(path, "rb");
    TBS_REQUIRE(fp, TBS_ERR_IO, "could not open %s", path);
    bool ok = read_n(fp, sizes, 3 * sizeof(int));
    fclose(fp);
    TBS_REQUIRE(ok, TBS_ERR_IO, "%s: truncated header", path);
    TBS_REQUIRE(sizes[0] >= 1 && sizes[1] >= 0 && sizes[2] >= 0, TBS_ERR_IO, "%s: bad header", path);
    TBS_REQUIRE(jac_file_bytes(sizes) == sz, TBS_ERR_IO,
                "%s: header implies %ld bytes, file has %ld", path, jac_file_bytes(sizes), sz);
    return TBS_OK;
}

int tbs_read_jacobian(const char *path, double *jacD, double *jacO1, double *jacO2,
                      double *rhs, double *U0) {
    int h[3];
    int rc = tbs_read_jacobian_header(path, h);
    if (rc) return rc;
    FILE *fp = fopen(path, "rb");
    TBS_REQUIRE(fp, TBS_ERR_IO, "could not open %s", path);
    int hh[3];
    bool ok = read_n(fp, hh, sizeof(hh));
    const long bs = (long)h[0] * h[0];
    struct { double *p; long n; } arr[5] = {{jacD, bs * h[1]}, {jacO1, bs * h[2]},
                                           {jacO2, bs * h[2]}, {rhs, (long)h[0] * h[1]},
                                           {U0, (long)h[0] * h[1]}};
    for (auto &a : arr) {
        if (!ok) break;
        if (a.p) ok = read_n(fp, a.p, a.n * sizeof(double));
        else ok = fseek(fp, a.n * (long)sizeof(double), SEEK_CUR) == 0;
    }
    fclose(fp);
    TBS_REQUIRE(ok, TBS_ERR_IO, "%s: short read", path);
    return TBS_OK;
}

int tbs_write_jacobian(const char *path, int nvar, int nelem, int nintface, const double *jacD,
                       const double *jacO1, const double *jacO2, const double *rhs,
                       const double *U0) {
    TBS_REQUIRE(path && jacD && jacO1 && jacO2 && rhs, TBS_ERR_ARG, "null argument");
    FILE *fp = fopen(path, "wb");
    TBS_REQUIRE(fp, TBS_ERR_IO, "could not create %s", path);
    int h[3] = {nvar, nelem, nintface};
    const long bs = (long)nvar * nvar;
    bool ok = fwrite(h, sizeof(int), 3, fp) == 3;
    ok = ok && fwrite(jacD, sizeof(double), bs * nelem, fp) == (size_t)(bs * nelem);
    ok = ok && fwrite(jacO1, sizeof(double), bs * nintface, fp) == (size_t)(bs * nintface);
    ok = ok && fwrite(jacO2, sizeof(double), bs * nintface, fp) == (size_t)(bs * nintface);
    ok = ok && fwrite(rhs, sizeof(double), (long)nvar * nelem, fp) == (size_t)((long)nvar * nelem);
    std::vector<double> zero;
    if (!U0) { zero.assign((size_t)nvar * nelem, 0.0); U0 = zero.data(); }
    ok = ok && fwrite(U0, sizeof(double), (long)nvar * nelem, fp) == (size_t)((long)nvar * nelem);
    ok = (fclose(fp) == 0) && ok;
    TBS_REQUIRE(ok, TBS_ERR_IO, "%s: write failed", path);
    return TBS_OK;
}

// Every index the drop-in kernels follow, checked on the host before any of
// them runs: the reference's kernels trust their mesh (src/okl/lineLU_v2.okl,
// src/okl/linesmoothLU_TriBlock_v3.okl index with it directly), so a
// corrupted file or a mesh of another system must be refused here.
int tbs_mesh_validate(const tbs_mesh_host *m) {
    TBS_REQUIRE(m != nullptr, TBS_ERR_ARG, "null mesh");
    TBS_REQUIRE(m->nvar >= 1 && m->nelem >= 1 && m->nintface >= 0 && m->eftot >= 0 && m->nline >= 0 &&
                    m->nlineelem >= 0 && m->nlineelem <= m->nelem,
                TBS_ERR_IO, "mesh sizes out of range (nvar %d nelem %d nintface %d eftot %d nline %d nlineelem %d)",
                m->nvar, m->nelem, m->nintface, m->eftot, m->nline, m->nlineelem);
    TBS_REQUIRE(m->epoint && (m->eftot == 0 || m->ef) && (m->nintface == 0 || m->fc) && m->linepoint &&
                    (m->nline == 0 || m->linesize) && (m->nlineelem == 0 || (m->lines && m->lineface)),
                TBS_ERR_ARG, "null mesh array");
    TBS_REQUIRE(m->epoint[0] == 0 && m->epoint[m->nelem] == m->eftot, TBS_ERR_IO,
                "epoint must run from 0 to eftot = %d", m->eftot);
    for (int e = 0; e < m->nelem; ++e)
        TBS_REQUIRE(m->epoint[e + 1] >= m->epoint[e], TBS_ERR_IO, "epoint decreases at element %d", e);
    for (int k = 0; k < m->eftot; ++k)
        TBS_REQUIRE(m->ef[k] < m->nintface, TBS_ERR_IO, "ef[%d] = %d is not a face (nintface %d)", k, m->ef[k],
                    m->nintface);
    for (long f = 0; f < 2L * m->nintface; ++f)
        TBS_REQUIRE(m->fc[f] >= 0 && m->fc[f] < m->nelem, TBS_ERR_IO, "fc[%ld] = %d is not an element", f, m->fc[f]);
    TBS_REQUIRE(m->linepoint[0] == 0, TBS_ERR_IO, "linepoint[0] must be 0");
    for (int l = 0; l < m->nline; ++l)
        TBS_REQUIRE(m->linesize[l] >= 1 && (long)m->linepoint[l] + m->linesize[l] == m->linepoint[l + 1],
                    TBS_ERR_IO, "line %d: linepoint/linesize not contiguous", l);
    TBS_REQUIRE(m->linepoint[m->nline] == m->nlineelem, TBS_ERR_IO, "linepoint[nline] != nlineelem");
    for (int i = 0; i < m->nlineelem; ++i)
        TBS_REQUIRE(m->lines[i] >= 0 && m->lines[i] < m->nelem, TBS_ERR_IO, "lines[%d] = %d is not an element",
                    i, m->lines[i]);
    for (int l = 0; l < m->nline; ++l)
        for (int k = 1; k < m->linesize[l]; ++k) {
            const int mm = m->linepoint[l] + k;
            const int f = m->lineface[mm];
            TBS_REQUIRE(f >= 0 && f < m->nintface, TBS_ERR_IO, "lineface[%d] = %d is not an interior face", mm, f);
            const int e = m->lines[mm], ep = m->lines[mm - 1];
            const int a = m->fc[2 * f], b = m->fc[2 * f + 1];
            TBS_REQUIRE((a == e && b == ep) || (a == ep && b == e), TBS_ERR_IO,
                        "lineface[%d] does not join consecutive line elements", mm);
        }
    return TBS_OK;
}

}  // extern "C"
