// lij_fast.hip -- the fused LIJ sweep path ("plan" / "batch"): host side and
// the kernels' instances for block sizes 2..16 (the kernels themselves are in
// lij_kernels.hpp; sizes 17..32 are instantiated in lij_wide_lo.hip and
// lij_wide_hi.hip).
#include "lij_kernels.hpp"
#include "lij_wide.hpp"

// ======================================================================
// Host side
// ======================================================================

struct tbs_plan {
    tbs_ctx *ctx = nullptr;
    int nv = 9, prec = TBS_FP64;
    int nelem = 0, nintface = 0, nline = 0, nle = 0, F = 0, S = 1, ngroups = 0, nwin = 0;
    bool factored = false;
    std::vector<int> h_linesize, h_linepoint, h_woff;
    // device: topology
    int *d_lines = nullptr, *d_mline = nullptr, *d_linepoint = nullptr, *d_linesize = nullptr;
    int *d_woff = nullptr;
    int *d_fc = nullptr, *d_lineface = nullptr;  // for lineLU / assemble
    int2 *d_src = nullptr;                       // packing source codes
    int4 *d_fidx = nullptr;                      // factor block sources per element
    // device: data
    double *d_blk = nullptr;
    int *d_nbr = nullptr;
    int *d_slotc = nullptr;
    void *d_W = nullptr, *d_EF = nullptr, *d_GB = nullptr;
    double *d_rhs = nullptr, *d_U = nullptr, *d_R = nullptr, *d_gpart = nullptr, *d_hist = nullptr;
    double *a_rhs = nullptr, *a_U = nullptr, *a_R = nullptr;   // allocations (kPad before element 0)
    void *d_g = nullptr, *d_dUp = nullptr;
    int *d_state = nullptr;
    // the plan as a one-system batch
    void *d_sys = nullptr;
    int4 *d_ltab = nullptr;
    int2 *d_gtab = nullptr;
    std::map<int, hipGraphExec_t> graphs;
    int max_sweeps = kHistCap - 1;
    double rtol = 0.0;
    int solves = 0;           // tbs_plan_solve calls so far
    int variant = 0;          // Thomas variant (autotuned): prefetch, LDS windows, lines per workgroup
    int rvariant = 0;         // residual workgroup shape (autotuned)
    double t_us[kTuneThomas] = {}, r_us[kTuneResidual] = {};   // last autotune's timings
    size_t bytes = 0;
    int *h_state = nullptr;      // pinned unless region-backed
    bool in_region = false;      // device arrays carved from a context region
    std::vector<void *> allocs;  // hipMalloc'd arrays (not region-backed)
    std::unique_ptr<char[]> h_tables;  // host image of the packed tables (source of an async upload)
};

struct tbs_batch {
    tbs_ctx *ctx = nullptr;
    std::vector<tbs_plan *> plans;
    int nv = 9, prec = TBS_FP64;
    void *d_sys = nullptr;
    int4 *d_ltab = nullptr;
    int2 *d_gtab = nullptr;
    int nl = 0, ng = 0;
    // eager sweeps may run in two halves on the context's two streams: systems
    // [0, nsA) with lines [0, nlA) and groups [0, ngA) of the tables, the
    // rest after them (nsA = 0: one system, one stream)
    int nsA = 0, nlA = 0, ngA = 0;
    std::map<std::pair<int, std::pair<int, double>>, hipGraphExec_t> graphs;
    int *h_states = nullptr;    // pinned [cap_sys][4]
    int *d_states = nullptr;    // [cap_sys][4] gathered system states
    int variant = 0, rvariant = 0;
    bool split = true;          // eager sweeps in two halves on two streams (autotuned)
    double t_us[kTuneThomas] = {}, r_us[kTuneResidual] = {}, pass_us[2] = {};
    // capacities of the device tables (entries) and of the pinned staging
    size_t cap_sys = 0, cap_l = 0, cap_g = 0, cap_stage = 0;
    char *h_stage = nullptr;    // pinned image of the tables being uploaded
    hipEvent_t staged = nullptr; // the last upload from h_stage has completed
};

namespace {

template <typename TF>
SysDev<TF> make_sysdev(const tbs_plan *p) {
    SysDev<TF> s{};
    s.blk = p->d_blk;
    s.nbr = p->d_nbr;
    s.slotc = p->d_slotc;
    s.W = (const TF *)p->d_W;
    s.EF = (const TF *)p->d_EF;
    s.GB = (const TF *)p->d_GB;
    s.rhs = p->d_rhs;
    s.U = p->d_U;
    s.g = (TF *)p->d_g;
    s.dUp = (TF *)p->d_dUp;
    s.R = p->d_R;
    s.gpart = p->d_gpart;
    s.hist = p->d_hist;
    s.state = p->d_state;
    s.n = p->nle;
    s.ngroups = p->ngroups;
    s.F = p->F;
    return s;
}

size_t tf_size(int prec) { return prec == TBS_FP32 ? sizeof(float) : sizeof(double); }

// Thomas factor-window stride in values (tbs::WinLayout::STRIDE for this nvar/precision).
size_t win_stride(int nv, int prec) {
    const size_t tf = tf_size(prec), lw = (size_t)(64 / nv) * nv;
    return (nv * lw * tf + 15) / 16 * 16 / tf;
}

int check_mesh(const tbs_mesh_host *m) {
    TBS_REQUIRE(m != nullptr, TBS_ERR_ARG, "null mesh");
    TBS_REQUIRE(m->nelem >= 1 && m->nline >= 1 && m->nintface >= 0 && m->eftot >= 0, TBS_ERR_ARG,
                "bad mesh sizes");
    TBS_REQUIRE(m->epoint && m->ef && m->linesize && m->linepoint && m->lines && m->lineface,
                TBS_ERR_ARG, "null mesh array");
    TBS_REQUIRE(m->nintface == 0 || m->fc, TBS_ERR_ARG, "null fc");
    TBS_REQUIRE(m->nlineelem == m->nelem, TBS_ERR_ARG,
                "the fused path needs every element in exactly one line (nlineelem %d != nelem %d)",
                m->nlineelem, m->nelem);
    TBS_REQUIRE(m->linepoint[0] == 0, TBS_ERR_ARG, "linepoint[0] must be 0");
    for (int l = 0; l < m->nline; ++l)
        TBS_REQUIRE(m->linesize[l] >= 1 && m->linepoint[l + 1] == m->linepoint[l] + m->linesize[l],
                    TBS_ERR_ARG, "line %d: linepoint/linesize not contiguous", l);
    TBS_REQUIRE(m->linepoint[m->nline] == m->nlineelem, TBS_ERR_ARG, "linepoint[nline] != nlineelem");
    std::vector<char> seen(m->nelem, 0);
    for (int i = 0; i < m->nlineelem; ++i) {
        const int e = m->lines[i];
        TBS_REQUIRE(e >= 0 && e < m->nelem && !seen[e], TBS_ERR_ARG, "lines[] is not a permutation");
        seen[e] = 1;
    }
    TBS_REQUIRE(m->epoint[0] == 0 && m->epoint[m->nelem] == m->eftot, TBS_ERR_ARG, "bad epoint");
    for (int e = 0; e < m->nelem; ++e)
        TBS_REQUIRE(m->epoint[e + 1] >= m->epoint[e], TBS_ERR_ARG, "epoint not monotone");
    for (int k = 0; k < m->eftot; ++k)
        TBS_REQUIRE(m->ef[k] < m->nintface, TBS_ERR_ARG, "ef[%d] out of range", k);
    for (int f = 0; f < 2 * m->nintface; ++f)
        TBS_REQUIRE(m->fc[f] >= 0 && m->fc[f] < m->nelem, TBS_ERR_ARG, "fc[%d] out of range", f);
    for (int l = 0; l < m->nline; ++l)
        for (int k = 1; k < m->linesize[l]; ++k) {
            const int mm = m->linepoint[l] + k;
            const int f = m->lineface[mm];
            TBS_REQUIRE(f >= 0 && f < m->nintface, TBS_ERR_ARG, "lineface[%d] not an interior face", mm);
            const int e = m->lines[mm], ep = m->lines[mm - 1];
            const int a = m->fc[2 * f], b = m->fc[2 * f + 1];
            TBS_REQUIRE((a == e && b == ep) || (a == ep && b == e), TBS_ERR_ARG,
                        "lineface[%d] does not join consecutive line elements", mm);
        }
    return TBS_OK;
}

// Sizes the plan's 32-bit buffer descriptors can address: whole-system
// vectors (< 4 GiB), and per line the factor windows and the vector values
// (< 2^31 bytes each, so the out-of-range base 2^31 stays outside them).
int check_ranges(const tbs_mesh_host *m, int precision) {
    TBS_REQUIRE(((size_t)m->nlineelem + 2 * kPad) * m->nvar * 8 < ((size_t)1 << 32) - 4096, TBS_ERR_UNSUPPORTED,
                "system too large for one plan: %d elements x %d values exceeds 4 GiB per vector",
                m->nlineelem, m->nvar);
    const size_t wbytes = win_stride(m->nvar, precision) * tf_size(precision);
    const int spw = 64 / m->nvar;
    for (int l = 0; l < m->nline; ++l) {
        const int n = m->linesize[l];
        TBS_REQUIRE((size_t)line_windows(n, spw) * wbytes < ((size_t)1 << 31) &&
                        (size_t)n * m->nvar * 8 < ((size_t)1 << 31),
                    TBS_ERR_UNSUPPORTED, "line %d has %d elements: its factors exceed 2 GiB, too long for one plan",
                    l, n);
    }
    return TBS_OK;
}

struct Layout {
    int F, S, ngroups, nwin;
    std::vector<int> woff;
};

// (F_known > 0: the faces per element are known without epoint, as for a
// structured grid whose tables the device generates)
Layout layout_of(const tbs_mesh_host *m, int F_known = 0) {
    Layout L;
    int F = F_known;
    for (int e = 0; F_known == 0 && e < m->nelem; ++e) F = std::max(F, m->epoint[e + 1] - m->epoint[e]);
    L.F = F;
    L.S = F + 1;
    L.ngroups = (m->nlineelem + kGroup - 1) / kGroup;
    const int spw = 64 / m->nvar;
    L.woff.resize(m->nline + 1);
    int w = 0;
    for (int l = 0; l < m->nline; ++l) {
        L.woff[l] = w;
        w += line_windows(m->linesize[l], spw);
    }
    L.woff[m->nline] = w;
    L.nwin = w;
    return L;
}

// Packed integer tables of a plan (one device array): byte offsets, each
// 16-byte aligned, and the total.
// The arrays up to host_end come from the host (mesh arrays as given, the
// window offsets and the system descriptor); the rest are built on the device.
struct TableLayout {
    size_t lines, linepoint, linesize, woff, fc, lineface, epoint, ef, sys, host_end;
    size_t mline, src, fidx, nbr, slotc, ltab, gtab, total;
};

TableLayout table_layout(const tbs_mesh_host *m, const Layout &L, int prec) {
    TableLayout t{};
    size_t o = 0;
    auto take = [&o](size_t bytes) { const size_t at = o; o += (bytes + 15) / 16 * 16; return at; };
    const size_t G = (size_t)L.ngroups * kGroup, Fm = (size_t)std::max(L.F, 1);
    const size_t nle = m->nlineelem, nl = m->nline;
    t.lines = take(4 * nle);
    t.linepoint = take(4 * (nl + 1));
    t.linesize = take(4 * nl);
    t.woff = take(4 * (nl + 1));
    t.fc = take(4 * std::max<size_t>(2 * (size_t)m->nintface, 1));
    t.lineface = take(4 * nle);
    t.epoint = take(4 * ((size_t)m->nelem + 1));
    t.ef = take(4 * std::max<size_t>((size_t)m->eftot, 1));
    t.sys = take(prec == TBS_FP32 ? sizeof(SysDev<float>) : sizeof(SysDev<double>));
    t.host_end = o;
    t.mline = take(4 * nle);
    t.src = take(sizeof(int2) * G * L.S);
    t.fidx = take(sizeof(int4) * nle);
    t.nbr = take(4 * G * Fm);
    t.slotc = take(4 * (size_t)L.ngroups * Fm);
    t.ltab = take(sizeof(int4) * nl);
    t.gtab = take(sizeof(int2) * L.ngroups);
    t.total = o;
    return t;
}

// Device arrays of a plan, in allocation order.
enum { A_BLK, A_W, A_EF, A_GB, A_RHS, A_U, A_R, A_G, A_DUP, A_GPART, A_HIST, A_STATE, A_TABLES, A_COUNT };

std::array<size_t, A_COUNT> array_sizes(const tbs_mesh_host *m, const Layout &L, int prec) {
    const size_t nv = m->nvar, bs = nv * nv, tf = tf_size(prec);
    const size_t G = (size_t)L.ngroups * kGroup, nle = m->nlineelem;
    std::array<size_t, A_COUNT> a{};
    a[A_BLK] = G * L.S * bs * 8;
    a[A_W] = G * bs * tf;
    a[A_EF] = a[A_GB] = (size_t)L.nwin * win_stride((int)nv, prec) * tf;
    a[A_RHS] = a[A_U] = a[A_R] = (nle + 2 * (size_t)kPad) * nv * 8;
    a[A_G] = a[A_DUP] = nle * nv * tf;
    a[A_GPART] = (size_t)L.ngroups * 8;
    a[A_HIST] = (size_t)kHistCap * 8;
    a[A_STATE] = 4 * sizeof(int);
    a[A_TABLES] = table_layout(m, L, prec).total;
    return a;
}

size_t aligned_bytes(size_t b) { return (std::max<size_t>(b, 1) + tbs::kRegionAlign - 1) / tbs::kRegionAlign * tbs::kRegionAlign; }

// Device bytes a plan holds (each array rounded to the region alignment).
size_t footprint_of(const tbs_mesh_host *m, const Layout &L, int prec) {
    size_t b = 0;
    for (size_t x : array_sizes(m, L, prec)) b += aligned_bytes(x);
    return b;
}

// Bytes of scratch the factorisation needs beyond the plan (none: the fused
// factor kernel writes the sweep-form factors directly).
size_t factor_workspace_bytes(int, int) { return 0; }

// Algorithmic bytes of one sweep, BASELINE model: factor-side D', A, DinvC
// (tf bytes each), D + Of1 + Of2 at FP64, 6 vector passes at FP64.
double model_bytes(const tbs_mesh_host *m, int prec) {
    const double bs = (double)m->nvar * m->nvar, nv = m->nvar;
    const double tf = (double)tf_size(prec);
    return tf * bs * (3.0 * m->nelem - 2.0 * m->nline) + 8.0 * bs * (m->nelem + 2.0 * m->nintface) +
           6.0 * 8.0 * nv * m->nelem;
}

// The reference's own per-sweep byte model, src/main.cxx:149-175 / iters.
double reference_model_bytes(const tbs_mesh_host *m) {
    const double nv = m->nvar, bs = nv * nv;
    const double du = 8.0 * (bs * m->nelem + bs * m->nelem + bs * m->nlineelem + 4.0 * nv * m->nelem +
                             nv * m->nelem) +
                      4.0 * (m->nline + (m->nline + 1.0) + 2.0 * m->nlineelem);
    const double cp = 5.0 * nv * m->nelem * 8.0;
    const double lr = 8.0 * (bs * m->nelem + 2.0 * bs * m->nintface + 5.0 * m->nelem * nv +
                             2.0 * nv * m->nelem) +
                      4.0 * ((m->nelem + 1.0) + m->eftot + 2.0 * m->nintface + m->nline +
                             (m->nline + 1.0) + 2.0 * m->nlineelem);
    return du + cp + lr;
}

int plan_alloc(tbs_plan *p, size_t bytes, void **ptr) {
    bool reg = false;
    TBS_TRY(tbs::dev_alloc(p->ctx, bytes, ptr, &reg));
    if (reg) p->in_region = true;
    else p->allocs.push_back(*ptr);
    p->bytes += aligned_bytes(bytes);
    return TBS_OK;
}

void free_plan(tbs_plan *p) {
    if (!p) return;
    hipSetDevice(p->ctx->device);
    for (auto &kv : p->graphs) hipGraphExecDestroy(kv.second);
    for (void *q : p->allocs) hipFree(q);
    if (p->h_state) {
        if (p->in_region) delete[] p->h_state;
        else hipHostFree(p->h_state);
    }
    delete p;
}

int thomas_variants(int nv, int prec) {
    if (nv == 9) return prec == TBS_FP32 ? ThomasV<9, float>::n : ThomasV<9, double>::n;
    return 1;
}

// encoded as 100 * PF + K + 10000 * (lines per workgroup - 1)
int thomas_pf_of(int nv, int prec, int variant) {
    TV t = {nv > 16 ? 1 : 2, 2, 1};
    if (nv == 9) t = prec == TBS_FP32 ? ThomasV<9, float>::v[variant % ThomasV<9, float>::n]
                                      : ThomasV<9, double>::v[variant % ThomasV<9, double>::n];
    return 100 * t.pf + t.k + 10000 * (t.lpw - 1);
}

// Dispatch on (nvar, precision) for the fused kernels: 2..16 instantiated
// here, 17..24 and 25..32 in lij_wide_lo.hip / lij_wide_hi.hip.
#define TBS_PLAN_NVARS(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
#define TBS_WIDE_CALL(nv, fn, ...)                                    \
    do {                                                              \
        if ((nv) >= 17 && (nv) <= 24) return tbs::wide_lo::fn(__VA_ARGS__); \
        if ((nv) >= 25 && (nv) <= 32) return tbs::wide_hi::fn(__VA_ARGS__); \
    } while (0)

FactorRec factor_rec(const tbs_plan *p, const double *D, const double *O1, const double *O2) {
    FactorRec r{};
    r.linepoint = p->d_linepoint;
    r.linesize = p->d_linesize;
    r.woff = p->d_woff;
    r.fidx = p->d_fidx;
    r.D = D;
    r.O1 = O1;
    r.O2 = O2;
    r.zero = p->a_rhs;   // the rhs allocation starts with kPad * nv >= nv * nv zeros
    r.W = p->d_W;
    r.EF = p->d_EF;
    r.GB = p->d_GB;
    return r;
}

int launch_factor_wide(int nv, int prec, hipStream_t st, const FactorRec &one, int nline, const FactorRec *recs,
                       const int2 *ltab) {
    TBS_WIDE_CALL(nv, factor, nv, prec, st, one, nline, recs, ltab);
    tbs::set_error("nvar %d not built for the fused path", nv);
    return TBS_ERR_UNSUPPORTED;
}

int launch_factor(int nv, int prec, hipStream_t st, const FactorRec &one, int nline, const FactorRec *recs,
                  const int2 *ltab) {
    if (nline == 0) return TBS_OK;
    if (nv > 16) TBS_TRY(launch_factor_wide(nv, prec, st, one, nline, recs, ltab));
    switch (nv) {
#define X(v) case v: \
        if (prec == TBS_FP32) launch_factor_nv<v, float>(st, one, nline, recs, ltab); \
        else launch_factor_nv<v, double>(st, one, nline, recs, ltab); \
        break;
        TBS_PLAN_NVARS(X)
#undef X
    }
    TBS_LAUNCHED("ffactor_kernel");
    return TBS_OK;
}

int launch_pack(const tbs_plan *p, hipStream_t pst, const double *jacD, const double *jacO1, const double *jacO2) {
    const int nv = p->nv, BS = nv * nv;
    const long nws = (long)p->ngroups * p->S;
    if (nv <= 16) {
        switch (nv) {
#define X(v) case v: launch_pack_lds<v>(pst, nws, p->S, p->d_src, jacD, jacO1, jacO2, p->d_blk); break;
            TBS_PLAN_NVARS(X)
#undef X
        }
        TBS_LAUNCHED("pack_lds_kernel");
    } else {
        const long total = (long)p->ngroups * kGroup * p->S * BS;
        if (nv > 16) {
            int rc = TBS_ERR_UNSUPPORTED;
            if (nv <= 24) rc = tbs::wide_lo::pack(nv, pst, total, p->S, p->d_src, jacD, jacO1, jacO2, p->d_blk);
            else rc = tbs::wide_hi::pack(nv, pst, total, p->S, p->d_src, jacD, jacO1, jacO2, p->d_blk);
            TBS_TRY(rc);
        }
        switch (nv) {
#define X(v) case v: hipLaunchKernelGGL(pack_blocks_kernel<v>, dim3(gs_grid(total)), dim3(256), 0, pst, \
                   total, p->S, p->d_src, jacD, jacO1, jacO2, p->d_blk); break;
            TBS_PLAN_NVARS(X)
#undef X
        }
        TBS_LAUNCHED("pack_blocks_kernel");
    }
    return TBS_OK;
}

int k_residual(int nv, int prec, hipStream_t st, const void *sys, const int2 *gtab, int ng,
               int want_R, int force, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, int rvariant = 0) {
    switch (nv) {
#define X(v) case v: return prec == TBS_FP32 ? launch_residual<v, float>(st, sys, gtab, ng, want_R, force, e0, e1, rvariant) \
                                         : launch_residual<v, double>(st, sys, gtab, ng, want_R, force, e0, e1, rvariant);
        TBS_PLAN_NVARS(X)
#undef X
    }
    TBS_WIDE_CALL(nv, residual, nv, prec, st, sys, gtab, ng, want_R, force, e0, e1);
    tbs::set_error("nvar %d not built for the fused path", nv);
    return TBS_ERR_UNSUPPORTED;
}
int k_thomas(int nv, int prec, hipStream_t st, const void *sys, const int4 *ltab, int nl, int variant,
             hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    switch (nv) {
#define X(v) case v: return prec == TBS_FP32 ? launch_thomas<v, float>(st, sys, ltab, nl, variant, e0, e1) \
                                         : launch_thomas<v, double>(st, sys, ltab, nl, variant, e0, e1);
        TBS_PLAN_NVARS(X)
#undef X
    }
    TBS_WIDE_CALL(nv, thomas, nv, prec, st, sys, ltab, nl, variant, e0, e1);
    tbs::set_error("nvar %d not built for the fused path", nv);
    return TBS_ERR_UNSUPPORTED;
}
int k_final(int prec, hipStream_t st, const void *sys, int nsys, int max_sweeps, double rtol, int force,
            hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    return prec == TBS_FP32 ? launch_final<float>(st, sys, nsys, max_sweeps, rtol, force, e0, e1)
                            : launch_final<double>(st, sys, nsys, max_sweeps, rtol, force, e0, e1);
}
int k_reset(int prec, hipStream_t st, const void *sys, int nsys) {
    if (prec == TBS_FP32)
        hipLaunchKernelGGL((reset_state_kernel<float>), dim3((nsys + 63) / 64), dim3(64), 0, st,
                           (const SysDev<float> *)sys, nsys);
    else
        hipLaunchKernelGGL((reset_state_kernel<double>), dim3((nsys + 63) / 64), dim3(64), 0, st,
                           (const SysDev<double> *)sys, nsys);
    TBS_LAUNCHED("reset_state_kernel");
    return TBS_OK;
}

bool plan_nvar_ok(int nv) { return nv >= 2 && nv <= 32; }

// One sweep: Thomas (+U update), residual, norm/convergence.
int batch_sweeps(tbs_batch *b, hipStream_t st, int nsweeps, int max_sweeps, double rtol, bool split);

// One sweep of the systems the tables ltab[0, nl) and gtab[0, ng) cover
// (they index sys by system); the norm step runs over sys_final[0, nsys).
int enqueue_sweep(int nv, int prec, hipStream_t st, const void *sys, const void *sys_final, int nsys,
                  const int4 *ltab, int nl, const int2 *gtab, int ng, int max_sweeps, double rtol, int variant,
                  int rvariant) {
    TBS_TRY(k_thomas(nv, prec, st, sys, ltab, nl, variant));
    TBS_TRY(k_residual(nv, prec, st, sys, gtab, ng, 0, 0, nullptr, nullptr, rvariant));
    TBS_TRY(k_final(prec, st, sys_final, nsys, max_sweeps, rtol, 0));
    return TBS_OK;
}

int capture_sweeps(hipStream_t st, int nsweeps, hipGraphExec_t *out,
                   const std::function<int()> &one_sweep) {
    TBS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = TBS_OK;
    for (int i = 0; i < nsweeps && rc == TBS_OK; ++i) rc = one_sweep();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) { if (graph) hipGraphDestroy(graph); return rc; }
    TBS_HIP(e);
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    TBS_HIP(e);
    return TBS_OK;
}

}  // namespace

namespace tbs {

int plan_scatter_U(tbs_plan *p, hipStream_t st, double *U) {
    const long total = (long)p->nle * p->nv;
    hipLaunchKernelGGL(scatter_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv, p->d_lines,
                       p->d_U, U);
    TBS_LAUNCHED("scatter U");
    return TBS_OK;
}

const double *plan_hist(const tbs_plan *p) { return p->d_hist; }

int plan_vecs(tbs_plan *p, PlanVecs *out) {
    TBS_REQUIRE(p && out && p->factored, TBS_ERR_STATE, "plan not factored");
    out->ctx = p->ctx;
    out->rhs = p->d_rhs;
    out->U = p->d_U;
    out->R = p->d_R;
    out->n = p->nle;
    out->nv = p->nv;
    out->lines = p->d_lines;
    return TBS_OK;
}

int plan_activate(tbs_plan *p, hipStream_t st) { return k_reset(p->prec, st, p->d_sys, 1); }

int plan_residual(tbs_plan *p, hipStream_t st, int want_R, bool u_zero) {
    return k_residual(p->nv, p->prec, st, p->d_sys, p->d_gtab, p->ngroups, want_R, u_zero ? 3 : 1, nullptr, nullptr,
                      p->rvariant);
}

int plan_thomas(tbs_plan *p, hipStream_t st) {
    return k_thomas(p->nv, p->prec, st, p->d_sys, p->d_ltab, p->nline, p->variant);
}

int plan_gather(tbs_plan *p, hipStream_t st, const double *src, double *dst) {
    const long total = (long)p->nle * p->nv;
    hipLaunchKernelGGL(gather_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv, p->d_lines, src,
                       dst);
    TBS_LAUNCHED("gather");
    return TBS_OK;
}

int plan_scatter(tbs_plan *p, hipStream_t st, const double *src, double *dst) {
    const long total = (long)p->nle * p->nv;
    hipLaunchKernelGGL(scatter_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv, p->d_lines,
                       src, dst);
    TBS_LAUNCHED("scatter");
    return TBS_OK;
}

size_t plan_factor_workspace(int nvar, int nelem) { return factor_workspace_bytes(nvar, nelem); }

}  // namespace tbs

extern "C" {

int tbs_plan_supported_nvar(int *out, int cap) {
    int n = 0;
    for (int v = 2; v <= 32; ++v, ++n)
        if (out && n < cap) out[n] = v;
    return n;
}

int tbs_plan_footprint(const tbs_mesh_host *mesh, int precision, size_t *bytes) {
    return tbs::plan_footprint(mesh, precision, bytes, true);
}

int tbs_plan_create(tbs_ctx *ctx, const tbs_mesh_host *m, int precision, tbs_plan **out) {
    return tbs::plan_create(ctx, m, precision, out, true);
}

}  // extern "C"

int tbs::plan_footprint(const tbs_mesh_host *mesh, int precision, size_t *bytes, bool validate, bool structured) {
    TBS_REQUIRE(bytes, TBS_ERR_ARG, "null output");
    if (validate) TBS_TRY(check_mesh(mesh));
    else TBS_REQUIRE(mesh != nullptr && mesh->nelem >= 1 && mesh->nline >= 1, TBS_ERR_ARG, "bad mesh sizes");
    TBS_REQUIRE(precision == TBS_FP64 || precision == TBS_FP32, TBS_ERR_ARG, "bad precision");
    TBS_REQUIRE(plan_nvar_ok(mesh->nvar), TBS_ERR_UNSUPPORTED, "nvar %d not built for the fused path", mesh->nvar);
    TBS_TRY(check_ranges(mesh, precision));
    Layout L = layout_of(mesh, structured ? 4 : 0);
    *bytes = footprint_of(mesh, L, precision);
    return TBS_OK;
}

// Plan creation; validate = false skips the O(mesh) consistency checks for
// meshes the library built itself (a served request's structured grid);
// structured_nk > 0: the mesh is tbs_structured_mesh's NL x NK grid, given
// by its sizes and line arrays only, and its element/face tables are
// generated on the device (no O(mesh) host work).
int tbs::plan_create(tbs_ctx *ctx, const tbs_mesh_host *m, int precision, tbs_plan **out, bool validate,
                     int structured_nk) {
    TBS_TRY(tbs::use_ctx(ctx));
    TBS_REQUIRE(out, TBS_ERR_ARG, "null output");
    *out = nullptr;
    if (validate) TBS_TRY(check_mesh(m));
    else TBS_REQUIRE(m != nullptr && m->nelem >= 1 && m->nline >= 1 && m->nlineelem == m->nelem, TBS_ERR_ARG,
                     "bad mesh sizes");
    TBS_REQUIRE(m->nvar == ctx->nvar, TBS_ERR_ARG, "mesh nvar %d != context nvar %d", m->nvar, ctx->nvar);
    TBS_REQUIRE(plan_nvar_ok(m->nvar), TBS_ERR_UNSUPPORTED, "nvar %d not built for the fused path", m->nvar);
    TBS_REQUIRE(precision == TBS_FP64 || precision == TBS_FP32, TBS_ERR_ARG, "bad precision");
    TBS_TRY(check_ranges(m, precision));
    const int nv = m->nvar;
    const bool gen = structured_nk > 0;
    TBS_REQUIRE(!gen || (m->nline >= 1 && (long)m->nline * structured_nk == m->nelem && m->nlineelem == m->nelem),
                TBS_ERR_ARG, "structured grid sizes do not match");
    Layout L = layout_of(m, gen ? 4 : 0);
    std::unique_ptr<tbs_plan, void (*)(tbs_plan *)> p(new tbs_plan(), free_plan);
    p->ctx = ctx;
    p->nv = nv;
    p->prec = precision;
    p->nelem = m->nelem;
    p->nintface = m->nintface;
    p->nline = m->nline;
    p->nle = m->nlineelem;
    p->F = L.F;
    p->S = L.S;
    p->ngroups = L.ngroups;
    p->nwin = L.nwin;
    p->h_linesize.assign(m->linesize, m->linesize + m->nline);
    p->h_linepoint.assign(m->linepoint, m->linepoint + m->nline + 1);
    p->h_woff = L.woff;
    const size_t G = (size_t)L.ngroups * kGroup;

    tbs_plan *P = p.get();
    const std::array<size_t, A_COUNT> sz = array_sizes(m, L, precision);
    void *arr[A_COUNT];
    for (int i = 0; i < A_COUNT; ++i) TBS_TRY(plan_alloc(P, sz[i], &arr[i]));
    P->d_blk = (double *)arr[A_BLK];
    P->d_W = arr[A_W];
    P->d_EF = arr[A_EF];
    P->d_GB = arr[A_GB];
    P->a_rhs = (double *)arr[A_RHS];
    P->a_U = (double *)arr[A_U];
    P->a_R = (double *)arr[A_R];
    P->d_rhs = P->a_rhs + (size_t)kPad * nv;
    P->d_U = P->a_U + (size_t)kPad * nv;
    P->d_R = P->a_R + (size_t)kPad * nv;
    P->d_g = arr[A_G];
    P->d_dUp = arr[A_DUP];
    P->d_gpart = (double *)arr[A_GPART];
    P->d_hist = (double *)arr[A_HIST];
    P->d_state = (int *)arr[A_STATE];
    char *tb = (char *)arr[A_TABLES];
    const TableLayout T = table_layout(m, L, precision);
    P->d_lines = (int *)(tb + T.lines);
    P->d_mline = (int *)(tb + T.mline);
    P->d_linepoint = (int *)(tb + T.linepoint);
    P->d_linesize = (int *)(tb + T.linesize);
    P->d_woff = (int *)(tb + T.woff);
    P->d_fc = (int *)(tb + T.fc);
    P->d_lineface = (int *)(tb + T.lineface);
    P->d_src = (int2 *)(tb + T.src);
    P->d_fidx = (int4 *)(tb + T.fidx);
    P->d_nbr = (int *)(tb + T.nbr);
    P->d_slotc = (int *)(tb + T.slotc);
    P->d_ltab = (int4 *)(tb + T.ltab);
    P->d_gtab = (int2 *)(tb + T.gtab);
    P->d_sys = tb + T.sys;
    if (P->in_region) P->h_state = new int[4]();
    else TBS_HIP(hipHostMalloc((void **)&P->h_state, 4 * sizeof(int), hipHostMallocDefault));

    // host part of the tables (the mesh arrays as given): small plans (a
    // served request's) go through the context's pinned staging buffer,
    // reused from plan to plan, so the upload is one DMA; large ones through
    // a pageable image kept by the plan so the upload stays asynchronous
    constexpr size_t kStageMax = (size_t)64 << 20;
    const bool staged = T.host_end <= kStageMax || gen;
    // (a generated grid stages only its line arrays and its system record:
    // the image holds [linepoint, fc) and then [sys, host_end), and put()
    // shifts the offsets accordingly)
    const size_t lo_end = T.fc - T.linepoint;
    const size_t img = gen ? lo_end + (T.host_end - T.sys) : T.host_end;
    char *h = nullptr;
    if (staged) {
        TBS_TRY(tbs::stage_acquire(ctx, img, &h));
    } else {
        P->h_tables.reset(new char[T.host_end]);
        h = P->h_tables.get();
    }
    const int nle = P->nle, F = L.F, Fm = std::max(F, 1);
    auto at = [&](size_t off) -> size_t {
        return !gen ? off : (off >= T.sys ? lo_end + (off - T.sys) : off - T.linepoint);
    };
    auto put = [&](size_t off, const void *src_, size_t bytes) { if (bytes) std::memcpy(h + at(off), src_, bytes); };
    if (!gen) put(T.lines, m->lines, (size_t)nle * 4);
    put(T.linepoint, m->linepoint, (size_t)(P->nline + 1) * 4);
    put(T.linesize, m->linesize, (size_t)P->nline * 4);
    put(T.woff, L.woff.data(), (size_t)(P->nline + 1) * 4);
    if (!gen) {
        put(T.fc, m->fc, (size_t)2 * m->nintface * 4);
        put(T.lineface, m->lineface, (size_t)nle * 4);
        put(T.epoint, m->epoint, (size_t)(P->nelem + 1) * 4);
        put(T.ef, m->ef, (size_t)m->eftot * 4);
    }
    if (precision == TBS_FP32) {
        SysDev<float> sd = make_sysdev<float>(P);
        put(T.sys, &sd, sizeof(sd));
    } else {
        SysDev<double> sd = make_sysdev<double>(P);
        put(T.sys, &sd, sizeof(sd));
    }
    hipStream_t st = ctx->stream;
    if (gen) {
        // the line arrays and the system record from the host; the element
        // and face tables from the generator kernel
        TBS_HIP(hipMemcpyAsync(tb + T.linepoint, h, lo_end, hipMemcpyHostToDevice, st));
        TBS_HIP(hipMemcpyAsync(tb + T.sys, h + lo_end, T.host_end - T.sys, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(structured_tables_kernel, dim3(gs_grid(nle)), dim3(256), 0, st, m->nline, structured_nk,
                           P->d_lines, P->d_fc, P->d_lineface, (int *)(tb + T.epoint), (int *)(tb + T.ef));
        TBS_LAUNCHED("structured_tables_kernel");
    } else {
        TBS_HIP(hipMemcpyAsync(tb, h, T.host_end, hipMemcpyHostToDevice, st));
    }
    if (staged) TBS_TRY(tbs::stage_release(ctx));
    // derived tables on the device; the R vector (zeroed below) holds the
    // inverse permutation meanwhile
    {
        const int *d_epoint = (const int *)(tb + T.epoint), *d_ef = (const int *)(tb + T.ef);
        int *inv = (int *)P->a_R;
        hipLaunchKernelGGL(tab_lines_kernel, dim3(P->nline), dim3(64), 0, st, P->d_linepoint, P->d_linesize,
                           P->d_woff, P->d_lines, P->d_lineface, P->d_fc, P->d_mline, P->d_fidx, P->d_ltab);
        TBS_LAUNCHED("tab_lines_kernel");
        hipLaunchKernelGGL(tab_inv_kernel, dim3(gs_grid(nle)), dim3(256), 0, st, nle, P->d_lines, inv);
        TBS_LAUNCHED("tab_inv_kernel");
        hipLaunchKernelGGL(tab_src_kernel, dim3(gs_grid((long)G)), dim3(256), 0, st, (long)G, nle, L.S, Fm,
                           P->d_lines, d_epoint, d_ef, P->d_fc, inv, P->d_src, P->d_nbr);
        TBS_LAUNCHED("tab_src_kernel");
        hipLaunchKernelGGL(tab_slot_kernel, dim3(gs_grid((long)L.ngroups * Fm)), dim3(256), 0, st, L.ngroups, F,
                           Fm, nle, P->d_nbr, P->d_slotc, P->d_gtab);
        TBS_LAUNCHED("tab_slot_kernel");
    }
    for (double *a : {P->a_rhs, P->a_U, P->a_R}) TBS_HIP(hipMemsetAsync(a, 0, sz[A_RHS], st));
    *out = p.release();
    return TBS_OK;
}

extern "C" {

int tbs_plan_destroy(tbs_plan *plan) {
    if (!plan) return TBS_OK;
    hipSetDevice(plan->ctx->device);
    // a region-backed plan frees no device memory; its owner retires it only
    // after the plan's work has completed
    if (!plan->in_region) hipStreamSynchronize(plan->ctx->stream);
    free_plan(plan);
    return TBS_OK;
}

int tbs_plan_info_get(tbs_plan *p, tbs_plan_info *info) {
    TBS_REQUIRE(p && info, TBS_ERR_ARG, "null argument");
    tbs_mesh_host mh{};
    mh.nvar = p->nv;
    mh.nelem = p->nelem;
    mh.nintface = p->nintface;
    mh.nline = p->nline;
    mh.nlineelem = p->nle;
    mh.eftot = 0;
    info->nvar = p->nv;
    info->nelem = p->nelem;
    info->nline = p->nline;
    info->max_faces = p->F;
    info->precision = p->prec;
    info->device_bytes = p->bytes;
    info->alg_bytes_per_sweep = model_bytes(&mh, p->prec);
    info->ref_bytes_per_sweep = 0.0;
    // per-kernel algorithmic bytes of the fused form, the sweep's model bytes
    // split between its two kernels (they add up to alg_bytes_per_sweep at
    // FP64):
    //   Thomas: E + Gamma factor blocks (tf), g in, U in + out (the forward
    //   intermediate d' it writes and reads back is the implementation's, not
    //   the algorithm's: 2 tf nv n more bytes moved)
    //   residual: D + Of1 + Of2 (FP64), W (tf), U, rhs in, g out
    const double bs = (double)p->nv * p->nv, nv = p->nv, tf = (double)tf_size(p->prec);
    const double n = p->nle;
    info->thomas_bytes = tf * bs * 2.0 * (n - p->nline) + nv * n * (tf + 16.0);
    info->residual_bytes = 8.0 * bs * (n + 2.0 * p->nintface) + tf * bs * n + nv * n * (16.0 + tf);
    return TBS_OK;
}

int tbs_plan_factor(tbs_plan *p, const double *jacD, const double *jacO1, const double *jacO2) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    return tbs_plans_factor(p->ctx, &p, 1, &jacD, &jacO1, &jacO2);
}

int tbs_plans_factor(tbs_ctx *ctx, tbs_plan *const *plans, int n, const double *const *jacD,
                     const double *const *jacO1, const double *const *jacO2) {
    TBS_REQUIRE(ctx && plans && n >= 1 && jacD && jacO1 && jacO2, TBS_ERR_ARG, "bad arguments");
    TBS_TRY(tbs::use_ctx(ctx));
    const int nv = plans[0] ? plans[0]->nv : 0, prec = plans[0] ? plans[0]->prec : 0;
    long nl = 0;
    for (int i = 0; i < n; ++i) {
        const tbs_plan *p = plans[i];
        TBS_REQUIRE(p && p->ctx == ctx, TBS_ERR_ARG, "plan %d is null or on another context", i);
        TBS_REQUIRE(p->nv == nv && p->prec == prec, TBS_ERR_ARG, "plan %d: block size or precision differs", i);
        TBS_REQUIRE(jacD[i] && (p->nintface == 0 || (jacO1[i] && jacO2[i])), TBS_ERR_ARG,
                    "plan %d: null Jacobian array", i);
        nl += p->nline;
    }
    TBS_REQUIRE(nl < INT_MAX, TBS_ERR_ARG, "too many lines");
    hipStream_t st = ctx->stream;
    // sweep-form factors W, E, Gamma straight from the blocks: the factor
    // kernel writes every window entry the sweep reads; of W only the last
    // group's slot has lanes past the last element, zeroed here
    for (int i = 0; i < n; ++i) {
        const tbs_plan *p = plans[i];
        const size_t slot = (size_t)kGroup * nv * nv * tf_size(prec);
        TBS_HIP(hipMemsetAsync((char *)p->d_W + (size_t)(p->ngroups - 1) * slot, 0, slot, st));
    }
    hipStream_t pst;
    TBS_TRY(tbs::aux_stream(ctx, &pst));
    // The packing starts beside the factor kernel.  (Running a first round of
    // resident factor workgroups alone and the packing beside the rest was
    // faster with an earlier, slower factor kernel; with this one it is no
    // faster at 1,200-1,600 lines of 1,280 and 8-11% slower for 64 systems
    // of short lines, measured.)
    TBS_HIP(hipEventRecord(ctx->ev_fork, st));
    TBS_HIP(hipStreamWaitEvent(pst, ctx->ev_fork, 0));
    if (n == 1) {
        const FactorRec fr = factor_rec(plans[0], jacD[0], jacO1[0], jacO2[0]);
        TBS_TRY(launch_factor(nv, prec, st, fr, (int)nl, nullptr, nullptr));
    } else {
        // one launch over every plan's lines, longest first; the records and
        // the line table go up through the staging ring into a device buffer
        // owned by the context (stream-ordered reuse)
        std::vector<std::pair<int, int2>> order;
        order.reserve(nl);
        for (int i = 0; i < n; ++i)
            for (int l = 0; l < plans[i]->nline; ++l) order.push_back({plans[i]->h_linesize[l], make_int2(i, l)});
        std::stable_sort(order.begin(), order.end(),
                         [](const std::pair<int, int2> &a, const std::pair<int, int2> &b) { return a.first > b.first; });
        const size_t rec_bytes = (size_t)n * sizeof(FactorRec), lt_off = (rec_bytes + 15) / 16 * 16;
        const size_t bytes = lt_off + (size_t)nl * sizeof(int2);
        char *h = nullptr;
        TBS_TRY(tbs::stage_acquire(ctx, bytes, &h));
        FactorRec *rec = (FactorRec *)h;
        for (int i = 0; i < n; ++i) rec[i] = factor_rec(plans[i], jacD[i], jacO1[i], jacO2[i]);
        int2 *lt = (int2 *)(h + lt_off);
        for (long b = 0; b < nl; ++b) lt[b] = order[b].second;
        TBS_TRY(tbs::dev_stage(ctx, bytes));
        TBS_HIP(hipMemcpyAsync(ctx->dstage, h, bytes, hipMemcpyHostToDevice, st));
        TBS_TRY(tbs::stage_release(ctx));
        const FactorRec *drec = (const FactorRec *)ctx->dstage;
        const int2 *dlt = (const int2 *)(ctx->dstage + lt_off);
        TBS_TRY(launch_factor(nv, prec, st, FactorRec{}, (int)nl, drec, dlt));
    }
    // residual blocks, packed on the context's second stream while the
    // factor kernel (two waves per line) runs on the main one: queued after
    // the factor kernel (its waves take their CUs first) and forked after the
    // memsets; the main stream waits for the packing before anything after
    // the call
    for (int i = 0; i < n; ++i) TBS_TRY(launch_pack(plans[i], pst, jacD[i], jacO1[i], jacO2[i]));
    TBS_HIP(hipEventRecord(ctx->ev_join, pst));
    TBS_HIP(hipStreamWaitEvent(st, ctx->ev_join, 0));
    for (int i = 0; i < n; ++i) plans[i]->factored = true;
    return TBS_OK;
}

int tbs_plan_begin(tbs_plan *p, const double *rhs, const double *U0) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    TBS_TRY(tbs::use_ctx(p->ctx));
    TBS_REQUIRE(p->factored, TBS_ERR_STATE, "tbs_plan_begin before tbs_plan_factor");
    TBS_REQUIRE(rhs, TBS_ERR_ARG, "null rhs");
    hipStream_t st = p->ctx->stream;
    const long total = (long)p->nle * p->nv;
    TBS_TRY(k_reset(p->prec, st, p->d_sys, 1));
    hipLaunchKernelGGL(gather_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                       p->d_lines, rhs, p->d_rhs);
    TBS_LAUNCHED("gather rhs");
    if (U0) {
        hipLaunchKernelGGL(gather_u0_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                           p->d_lines, U0, p->d_U, p->d_state + 3);
    } else {
        hipLaunchKernelGGL(gather_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                           p->d_lines, U0, p->d_U);
    }
    TBS_LAUNCHED("gather U0");
    // R_0 = rhs + J U_0 (= rhs when U_0 = 0, the reference's start: then J U
    // is skipped, also for a given U_0 whose values are all zero) and ||R_0||
    TBS_TRY(k_residual(p->nv, p->prec, st, p->d_sys, p->d_gtab, p->ngroups, 0, U0 ? 1 | 4 : 3, nullptr, nullptr,
                       p->rvariant));
    TBS_TRY(k_final(p->prec, st, p->d_sys, 1, p->max_sweeps, p->rtol, 1));
    return TBS_OK;
}

int tbs_plan_launch_thomas(tbs_plan *p) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    TBS_TRY(tbs::use_ctx(p->ctx));
    return k_thomas(p->nv, p->prec, p->ctx->stream, p->d_sys, p->d_ltab, p->nline, p->variant);
}

int tbs_plan_launch_residual(tbs_plan *p) {
    TBS_TRY(tbs_plan_launch_kernel(p, TBS_STEP_RESIDUAL, nullptr, nullptr));
    return tbs_plan_launch_kernel(p, TBS_STEP_NORM, nullptr, nullptr);
}

int tbs_plan_launch_kernel(tbs_plan *p, int step, void *ev_start, void *ev_stop) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    TBS_REQUIRE((ev_start == nullptr) == (ev_stop == nullptr), TBS_ERR_ARG, "give both events or neither");
    TBS_TRY(tbs::use_ctx(p->ctx));
    hipStream_t st = p->ctx->stream;
    hipEvent_t e0 = (hipEvent_t)ev_start, e1 = (hipEvent_t)ev_stop;
    switch (step) {
        case TBS_STEP_THOMAS: return k_thomas(p->nv, p->prec, st, p->d_sys, p->d_ltab, p->nline, p->variant, e0, e1);
        case TBS_STEP_RESIDUAL:
            return k_residual(p->nv, p->prec, st, p->d_sys, p->d_gtab, p->ngroups, 0, 0, e0, e1, p->rvariant);
        case TBS_STEP_NORM: return k_final(p->prec, st, p->d_sys, 1, p->max_sweeps, p->rtol, 0, e0, e1);
    }
    tbs::set_error("unknown sweep step %d", step);
    return TBS_ERR_ARG;
}

int tbs_plan_launch_sweeps(tbs_plan *p, int nsweeps, int use_graph) {
    TBS_REQUIRE(p && nsweeps >= 0, TBS_ERR_ARG, "bad arguments");
    TBS_TRY(tbs::use_ctx(p->ctx));
    hipStream_t st = p->ctx->stream;
    auto one = [p, st]() {
        return enqueue_sweep(p->nv, p->prec, st, p->d_sys, p->d_sys, 1, p->d_ltab, p->nline, p->d_gtab,
                             p->ngroups, p->max_sweeps, p->rtol, p->variant, p->rvariant);
    };
    if (!use_graph) {
        for (int i = 0; i < nsweeps; ++i) TBS_TRY(one());
        return TBS_OK;
    }
    if (nsweeps == 0) return TBS_OK;
    auto it = p->graphs.find(nsweeps);
    if (it == p->graphs.end()) {
        hipGraphExec_t ge = nullptr;
        TBS_TRY(capture_sweeps(st, nsweeps, &ge, one));
        it = p->graphs.emplace(nsweeps, ge).first;
    }
    TBS_HIP(hipGraphLaunch(it->second, st));
    return TBS_OK;
}

int tbs_plan_end(tbs_plan *p, double *U, double *R_final) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    TBS_TRY(tbs::use_ctx(p->ctx));
    hipStream_t st = p->ctx->stream;
    const long total = (long)p->nle * p->nv;
    if (U) {
        hipLaunchKernelGGL(scatter_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                           p->d_lines, p->d_U, U);
        TBS_LAUNCHED("scatter U");
    }
    if (R_final) {
        TBS_TRY(k_residual(p->nv, p->prec, st, p->d_sys, p->d_gtab, p->ngroups, 1, 1, nullptr, nullptr, p->rvariant));
        hipLaunchKernelGGL(scatter_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                           p->d_lines, p->d_R, R_final);
        TBS_LAUNCHED("scatter R");
    }
    return TBS_OK;
}

int tbs_plan_history(tbs_plan *p, double *hist, int cap, int *count) {
    TBS_REQUIRE(p && count, TBS_ERR_ARG, "null argument");
    TBS_TRY(tbs::use_ctx(p->ctx));
    hipStream_t st = p->ctx->stream;
    TBS_HIP(hipMemcpyAsync(p->h_state, p->d_state, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
    TBS_HIP(hipStreamSynchronize(st));
    int k = std::min(p->h_state[0], kHistCap);
    *count = k;
    if (hist && cap > 0 && k > 0) {
        TBS_HIP(hipMemcpy(hist, p->d_hist, std::min(k, cap) * sizeof(double), hipMemcpyDeviceToHost));
    }
    return TBS_OK;
}

int tbs_plan_internal(tbs_plan *p, int which, void **dptr, size_t *bytes) {
    TBS_REQUIRE(p && dptr && bytes, TBS_ERR_ARG, "null argument");
    const size_t tf = tf_size(p->prec), nv = p->nv, n = p->nle, G = (size_t)p->ngroups * kGroup;
    switch (which) {
        case 0: *dptr = p->d_g; *bytes = n * nv * tf; break;
        case 1: *dptr = p->d_dUp; *bytes = n * nv * tf; break;
        case 2: *dptr = p->d_U; *bytes = n * nv * 8; break;
        case 3: *dptr = p->d_rhs; *bytes = n * nv * 8; break;
        case 4: *dptr = p->d_R; *bytes = n * nv * 8; break;
        case 5: *dptr = p->d_W; *bytes = G * nv * nv * tf; break;
        case 6: *dptr = p->d_EF; *bytes = (size_t)p->nwin * win_stride((int)nv, p->prec) * tf; break;
        case 7: *dptr = p->d_GB; *bytes = (size_t)p->nwin * win_stride((int)nv, p->prec) * tf; break;
        case 8: *dptr = p->d_blk; *bytes = G * p->S * nv * nv * 8; break;
        default: TBS_REQUIRE(false, TBS_ERR_ARG, "unknown internal array %d", which);
    }
    return TBS_OK;
}

}  // extern "C"

namespace {
int residual_variants(int nv) { return nv == 9 ? ResidualV<9>::n : 1; }
int residual_rw_of(int nv, int rvariant) { return nv == 9 ? ResidualV<9>::rw[rvariant % ResidualV<9>::n] : 4; }

// Mean microseconds of one call of fn over reps calls on st (after one
// warm-up call), from two timing events.
int time_calls(hipStream_t st, int reps, const std::function<int()> &fn, double *us) {
    TBS_TRY(fn());
    hipEvent_t e0, e1;
    TBS_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));   // timing only
    TBS_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
    int rc = TBS_OK;
    if (hipEventRecord(e0, st)) rc = TBS_ERR_STATE;
    for (int r = 0; r < reps && rc == TBS_OK; ++r) rc = fn();
    if (rc == TBS_OK && (hipEventRecord(e1, st) || hipEventSynchronize(e1))) rc = TBS_ERR_STATE;
    float ms = 0.f;
    if (rc == TBS_OK) hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc == TBS_ERR_STATE) tbs::set_error("timing events failed");
    *us = 1e3 * ms / reps;
    return rc;
}

// Time every compiled Thomas variant and residual shape on zeroed vectors
// (the kernels' cost does not depend on the values) over the systems the
// tables cover, and return the fastest of each.  Leaves the systems' state
// reset; callers start a solve with begin() afterwards.
int autotune_kernels(tbs_ctx *ctx, int nv, int prec, const void *sys, int nsys, const int4 *ltab, int nl,
                     const int2 *gtab, int ng, const std::vector<std::pair<double *, long>> &zero_vecs, int reps,
                     int *tbest, double *t_us, int *rbest, double *r_us) {
    hipStream_t st = ctx->stream;
    for (auto &zv : zero_vecs) {
        hipLaunchKernelGGL(zero_kernel, dim3(gs_grid(zv.second)), dim3(256), 0, st, zv.second, zv.first);
        TBS_LAUNCHED("zero_kernel");
    }
    TBS_TRY(k_reset(prec, st, sys, nsys));
    double best_t = 1e300;
    *tbest = 0;
    for (int v = 0; v < thomas_variants(nv, prec) && v < kTuneThomas; ++v) {
        double us = 0.0;
        TBS_TRY(time_calls(st, reps, [&]() { return k_thomas(nv, prec, st, sys, ltab, nl, v); }, &us));
        t_us[v] = us;
        if (us < best_t) { best_t = us; *tbest = v; }
    }
    best_t = 1e300;
    *rbest = 0;
    for (int v = 0; v < residual_variants(nv) && v < kTuneResidual; ++v) {
        double us = 0.0;
        TBS_TRY(time_calls(st, reps, [&]() { return k_residual(nv, prec, st, sys, gtab, ng, 0, 1, nullptr, nullptr, v); },
                           &us));
        r_us[v] = us;
        if (us < best_t) { best_t = us; *rbest = v; }
    }
    TBS_TRY(k_reset(prec, st, sys, nsys));
    TBS_HIP(hipStreamSynchronize(st));
    return TBS_OK;
}

// nsweeps sweeps of a batch, enqueued on st.  With split, a batch in two
// halves forks the second half onto the context's second stream once and
// joins it back after the last sweep, so the halves drift apart within the
// run: each half's kernels follow each other, and every system's arithmetic
// is the same as on one stream.  (Captured into a graph, the two branches
// made a solve slower, so graphs are captured on one stream.)
int batch_sweeps(tbs_batch *b, hipStream_t st, int nsweeps, int max_sweeps, double rtol, bool split) {
    const int ns = (int)b->plans.size();
    if (ns == 0 || nsweeps == 0) return TBS_OK;
    if (!split || b->nsA <= 0 || b->nsA >= ns) {
        for (int i = 0; i < nsweeps; ++i)
            TBS_TRY(enqueue_sweep(b->nv, b->prec, st, b->d_sys, b->d_sys, ns, b->d_ltab, b->nl, b->d_gtab, b->ng,
                                  max_sweeps, rtol, b->variant, b->rvariant));
        return TBS_OK;
    }
    hipStream_t aux;
    TBS_TRY(tbs::aux_stream(b->ctx, &aux));
    const size_t ssz = b->prec == TBS_FP32 ? sizeof(SysDev<float>) : sizeof(SysDev<double>);
    const void *sysB = (const char *)b->d_sys + ssz * (size_t)b->nsA;
    TBS_HIP(hipEventRecord(b->ctx->ev_fork, st));
    TBS_HIP(hipStreamWaitEvent(aux, b->ctx->ev_fork, 0));
    int rc = TBS_OK;
    for (int i = 0; i < nsweeps && rc == TBS_OK; ++i) {
        rc = enqueue_sweep(b->nv, b->prec, st, b->d_sys, b->d_sys, b->nsA, b->d_ltab, b->nlA, b->d_gtab, b->ngA,
                           max_sweeps, rtol, b->variant, b->rvariant);
        if (rc == TBS_OK)
            rc = enqueue_sweep(b->nv, b->prec, aux, b->d_sys, sysB, ns - b->nsA, b->d_ltab + b->nlA, b->nl - b->nlA,
                               b->d_gtab + b->ngA, b->ng - b->ngA, max_sweeps, rtol, b->variant, b->rvariant);
    }
    // join even after a failed launch, so that a capture on st can end
    TBS_HIP(hipEventRecord(b->ctx->ev_join, aux));
    TBS_HIP(hipStreamWaitEvent(st, b->ctx->ev_join, 0));
    return rc;
}
// reps Thomas + residual passes of a batch (no norm step: the systems stay
// active), on one stream or as the two halves of batch_sweeps, for timing.
int batch_pass(tbs_batch *b, hipStream_t st, int reps, bool split) {
    const int ns = (int)b->plans.size();
    auto pass = [b](hipStream_t s2, const int4 *lt, int nl, const int2 *gt, int ng) {
        TBS_TRY(k_thomas(b->nv, b->prec, s2, b->d_sys, lt, nl, b->variant));
        return k_residual(b->nv, b->prec, s2, b->d_sys, gt, ng, 0, 1, nullptr, nullptr, b->rvariant);
    };
    if (!split || b->nsA <= 0 || b->nsA >= ns) {
        for (int r = 0; r < reps; ++r) TBS_TRY(pass(st, b->d_ltab, b->nl, b->d_gtab, b->ng));
        return TBS_OK;
    }
    hipStream_t aux;
    TBS_TRY(tbs::aux_stream(b->ctx, &aux));
    TBS_HIP(hipEventRecord(b->ctx->ev_fork, st));
    TBS_HIP(hipStreamWaitEvent(aux, b->ctx->ev_fork, 0));
    int rc = TBS_OK;
    for (int r = 0; r < reps && rc == TBS_OK; ++r) {
        rc = pass(st, b->d_ltab, b->nlA, b->d_gtab, b->ngA);
        if (rc == TBS_OK) rc = pass(aux, b->d_ltab + b->nlA, b->nl - b->nlA, b->d_gtab + b->ngA, b->ng - b->ngA);
    }
    TBS_HIP(hipEventRecord(b->ctx->ev_join, aux));
    TBS_HIP(hipStreamWaitEvent(st, b->ctx->ev_join, 0));
    return rc;
}

void fill_tuning(int nv, int prec, int variant, int rvariant, int split, const double *t_us, const double *r_us,
                 const double *pass_us, tbs_tuning *o) {
    *o = tbs_tuning{};
    const int code = thomas_pf_of(nv, prec, variant);
    o->thomas_variant = variant;
    o->thomas_prefetch = (code % 10000) / 100;
    o->thomas_lds_windows = code % 100;
    o->thomas_lines_per_wg = code / 10000 + 1;
    o->residual_variant = rvariant;
    o->residual_waves_per_wg = residual_rw_of(nv, rvariant);
    o->two_stream = split;
    o->n_thomas = std::min(thomas_variants(nv, prec), kTuneThomas);
    o->n_residual = std::min(residual_variants(nv), kTuneResidual);
    for (int v = 0; v < kTuneThomas; ++v) o->thomas_us[v] = t_us[v];
    for (int v = 0; v < kTuneResidual; ++v) o->residual_us[v] = r_us[v];
    for (int v = 0; v < o->n_residual; ++v) o->residual_shapes[v] = residual_rw_of(nv, v);
    if (pass_us) { o->pass_us[0] = pass_us[0]; o->pass_us[1] = pass_us[1]; }
}
}  // namespace

extern "C" {

int tbs_plan_autotune(tbs_plan *p, int reps, int *chosen, double *us, int cap) {
    TBS_REQUIRE(p && reps >= 1, TBS_ERR_ARG, "bad arguments");
    TBS_TRY(tbs::use_ctx(p->ctx));
    TBS_REQUIRE(p->factored, TBS_ERR_STATE, "autotune before factor");
    const long nvec = (long)p->nle * p->nv;
    const size_t tf = tf_size(p->prec);
    std::vector<std::pair<double *, long>> zv = {{p->d_U, nvec}, {p->d_rhs, nvec},
                                                 {(double *)p->d_g, (long)(nvec * tf / 8)}};
    int tb = 0, rb = 0;
    TBS_TRY(autotune_kernels(p->ctx, p->nv, p->prec, p->d_sys, 1, p->d_ltab, p->nline, p->d_gtab, p->ngroups, zv,
                             reps, &tb, p->t_us, &rb, p->r_us));
    if (tb != p->variant || rb != p->rvariant) {
        for (auto &kv : p->graphs) hipGraphExecDestroy(kv.second);
        p->graphs.clear();
        p->variant = tb;
        p->rvariant = rb;
    }
    if (chosen) *chosen = tb;
    for (int v = 0; us && v < cap && v < kTuneThomas; ++v) us[v] = p->t_us[v];
    return TBS_OK;
}

int tbs_plan_tuning(tbs_plan *p, tbs_tuning *out) {
    TBS_REQUIRE(p && out, TBS_ERR_ARG, "null argument");
    fill_tuning(p->nv, p->prec, p->variant, p->rvariant, 0, p->t_us, p->r_us, nullptr, out);
    return TBS_OK;
}

int tbs_plan_set_tuning(tbs_plan *p, const tbs_tuning *t) {
    TBS_REQUIRE(p && t, TBS_ERR_ARG, "null argument");
    TBS_REQUIRE(t->thomas_variant >= 0 && t->thomas_variant < thomas_variants(p->nv, p->prec), TBS_ERR_ARG,
                "Thomas variant %d not built", t->thomas_variant);
    TBS_REQUIRE(t->residual_variant >= 0 && t->residual_variant < residual_variants(p->nv), TBS_ERR_ARG,
                "residual variant %d not built", t->residual_variant);
    if (t->thomas_variant != p->variant || t->residual_variant != p->rvariant) {
        for (auto &kv : p->graphs) hipGraphExecDestroy(kv.second);
        p->graphs.clear();
        p->variant = t->thomas_variant;
        p->rvariant = t->residual_variant;
    }
    return TBS_OK;
}

int tbs_plan_set_variant(tbs_plan *p, int variant) {
    TBS_REQUIRE(p && variant >= 0 && variant < thomas_variants(p->nv, p->prec), TBS_ERR_ARG,
                "variant %d not built", variant);
    if (variant != p->variant) {
        for (auto &kv : p->graphs) hipGraphExecDestroy(kv.second);
        p->graphs.clear();
        p->variant = variant;
    }
    return TBS_OK;
}

int tbs_plan_variants(tbs_plan *p, int *pf, int cap) {
    TBS_REQUIRE(p, TBS_ERR_ARG, "null plan");
    const int n = thomas_variants(p->nv, p->prec);
    for (int v = 0; v < n && v < cap; ++v) if (pf) pf[v] = thomas_pf_of(p->nv, p->prec, v);
    return n;
}

}  // extern "C"

extern "C" {

int tbs_plan_solve(tbs_plan *p, const double *rhs, double *U, int max_sweeps, double rtol,
                   double *hist, int *nsweeps, double *R_final) {
    TBS_REQUIRE(p && U, TBS_ERR_ARG, "null argument");
    TBS_REQUIRE(max_sweeps >= 0 && max_sweeps < kHistCap, TBS_ERR_ARG,
                "max_sweeps %d outside [0, %d)", max_sweeps, kHistCap);
    TBS_REQUIRE(rtol >= 0.0, TBS_ERR_ARG, "rtol < 0");
    TBS_TRY(tbs::use_ctx(p->ctx));
    if (p->max_sweeps != max_sweeps || p->rtol != rtol) {
        for (auto &kv : p->graphs) hipGraphExecDestroy(kv.second);
        p->graphs.clear();
        p->max_sweeps = max_sweeps;
        p->rtol = rtol;
    }
    TBS_TRY(tbs_plan_begin(p, rhs, U));
    hipStream_t st = p->ctx->stream;
    int done = 0;
    const int chunk = 4;
    // a plan's first solve launches its sweeps one kernel at a time (a graph
    // capture costs more than the launches it saves over a few sweeps);
    // later solves replay captured graphs
    const int use_graph = p->solves++ > 0 ? 1 : 0;
    while (done < max_sweeps) {
        const int k = std::min(chunk, max_sweeps - done);
        TBS_TRY(tbs_plan_launch_sweeps(p, k, use_graph));
        done += k;
        TBS_HIP(hipMemcpyAsync(p->h_state, p->d_state, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
        TBS_HIP(hipStreamSynchronize(st));
        if (!p->h_state[1]) break;
    }
    TBS_TRY(tbs_plan_end(p, U, R_final));
    int count = 0;
    TBS_TRY(tbs_plan_history(p, hist, max_sweeps + 1, &count));
    if (nsweeps) *nsweeps = count > 0 ? count - 1 : 0;
    if (p->h_state[2] == 2 || p->h_state[2] == 4) {
        tbs::set_error(p->h_state[2] == 2 ? "non-finite residual" : "residual diverging");
        return TBS_ERR_NUMERIC;
    }
    return TBS_OK;
}

// ------------------------------------------------------------ batch --
int tbs_batch_create(tbs_ctx *ctx, tbs_plan *const *plans, int nplans, tbs_batch **out) {
    TBS_TRY(tbs::use_ctx(ctx));
    TBS_REQUIRE(out && plans && nplans >= 1 && plans[0], TBS_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    std::unique_ptr<tbs_batch, int (*)(tbs_batch *)> b(new tbs_batch(), tbs_batch_destroy);
    b->ctx = ctx;
    b->nv = plans[0]->nv;
    b->prec = plans[0]->prec;
    b->split = b->prec == TBS_FP64;   // until autotuned: FP64 halves overlap, FP32 ones gain nothing
    TBS_HIP(hipEventCreateWithFlags(&b->staged, hipEventDisableTiming));
    TBS_TRY(tbs_batch_set_plans(b.get(), plans, nplans));
    *out = b.release();
    return TBS_OK;
}

}  // extern "C"

namespace {
// Device tables and pinned staging of a batch, grown (after the work that
// reads them has finished) to hold ns systems, nl lines and ng groups.
int batch_grow(tbs_batch *b, size_t ns, size_t nl, size_t ng) {
    hipStream_t st = b->ctx->stream;
    const size_t ssz = b->prec == TBS_FP32 ? sizeof(SysDev<float>) : sizeof(SysDev<double>);
    if (ns > b->cap_sys || nl > b->cap_l || ng > b->cap_g) {
        TBS_HIP(hipStreamSynchronize(st));
        if (ns > b->cap_sys) {
            const size_t c = std::max<size_t>(2 * ns, 16);
            if (b->d_sys) TBS_HIP(hipFree(b->d_sys));
            if (b->d_states) TBS_HIP(hipFree(b->d_states));
            if (b->h_states) TBS_HIP(hipHostFree(b->h_states));
            b->d_sys = nullptr; b->d_states = nullptr; b->h_states = nullptr; b->cap_sys = 0;
            TBS_HIP(hipMalloc(&b->d_sys, c * ssz));
            TBS_HIP(hipMalloc((void **)&b->d_states, c * 4 * sizeof(int)));
            TBS_HIP(hipHostMalloc((void **)&b->h_states, c * 4 * sizeof(int), hipHostMallocDefault));
            b->cap_sys = c;
        }
        if (nl > b->cap_l) {
            const size_t c = std::max<size_t>(2 * nl, 256);
            if (b->d_ltab) TBS_HIP(hipFree(b->d_ltab));
            b->d_ltab = nullptr; b->cap_l = 0;
            TBS_HIP(hipMalloc((void **)&b->d_ltab, c * sizeof(int4)));
            b->cap_l = c;
        }
        if (ng > b->cap_g) {
            const size_t c = std::max<size_t>(2 * ng, 256);
            if (b->d_gtab) TBS_HIP(hipFree(b->d_gtab));
            b->d_gtab = nullptr; b->cap_g = 0;
            TBS_HIP(hipMalloc((void **)&b->d_gtab, c * sizeof(int2)));
            b->cap_g = c;
        }
    }
    const size_t total = (ns * ssz + 15) / 16 * 16 + nl * sizeof(int4) + ng * sizeof(int2);
    if (total > b->cap_stage) {
        TBS_HIP(hipEventSynchronize(b->staged));   // the previous upload must have left it
        if (b->h_stage) TBS_HIP(hipHostFree(b->h_stage));
        b->h_stage = nullptr; b->cap_stage = 0;
        const size_t c = std::max<size_t>(2 * total, 4096);
        TBS_HIP(hipHostMalloc((void **)&b->h_stage, c, hipHostMallocDefault));
        b->cap_stage = c;
    }
    return TBS_OK;
}
}  // namespace

int tbs::batch_create_reserved(tbs_ctx *ctx, int nvar, int prec, size_t nsys, size_t nlines, size_t ngroups,
                               tbs_batch **out) {
    TBS_TRY(tbs::use_ctx(ctx));
    TBS_REQUIRE(out && (prec == TBS_FP64 || prec == TBS_FP32), TBS_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    std::unique_ptr<tbs_batch, int (*)(tbs_batch *)> b(new tbs_batch(), tbs_batch_destroy);
    b->ctx = ctx;
    b->nv = nvar;
    b->prec = prec;
    b->split = prec == TBS_FP64;
    TBS_HIP(hipEventCreateWithFlags(&b->staged, hipEventDisableTiming));
    TBS_HIP(hipEventRecord(b->staged, ctx->stream));
    TBS_TRY(batch_grow(b.get(), nsys, nlines, ngroups));
    *out = b.release();
    return TBS_OK;
}

extern "C" {

int tbs_batch_set_plans(tbs_batch *b, tbs_plan *const *plans, int nplans) {
    TBS_REQUIRE(b && nplans >= 0 && (plans || nplans == 0), TBS_ERR_ARG, "bad batch arguments");
    TBS_TRY(tbs::use_ctx(b->ctx));
    for (int i = 0; i < nplans; ++i) {
        TBS_REQUIRE(plans[i] && plans[i]->ctx->device == b->ctx->device, TBS_ERR_ARG,
                    "plan %d is on another device", i);
        TBS_REQUIRE(plans[i]->nv == b->nv && plans[i]->prec == b->prec, TBS_ERR_ARG,
                    "plans in a batch must share nvar and precision");
        TBS_REQUIRE(plans[i]->factored, TBS_ERR_STATE, "plan %d not factored", i);
    }
    hipStream_t st = b->ctx->stream;
    if (!b->graphs.empty()) {
        TBS_HIP(hipStreamSynchronize(st));
        for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
        b->graphs.clear();
    }
    // the systems in two halves of about equal element count, which eager
    // sweeps run on two streams when b->split is set (each half's Thomas pass
    // overlaps the other's residual)
    int nsA = 0;
    if (nplans >= 2) {
        long tot = 0, acc = 0;
        for (int s2 = 0; s2 < nplans; ++s2) tot += plans[s2]->nle;
        while (nsA < nplans - 1 && 2 * (acc + plans[nsA]->nle) <= tot) acc += plans[nsA++]->nle;
        nsA = std::max(nsA, 1);
    }
    // tables: each half's lines longest-first so the longest chains start
    // first; groups in system order
    std::vector<std::pair<int, int4>> lines;
    std::vector<int2> gtab;
    size_t nlA = 0, ngA = 0;
    auto longest_first = [](const std::pair<int, int4> &x, const std::pair<int, int4> &y) { return x.first > y.first; };
    for (int s2 = 0; s2 < nplans; ++s2) {
        if (s2 == nsA && nsA > 0) {
            std::stable_sort(lines.begin(), lines.end(), longest_first);
            nlA = lines.size();
            ngA = gtab.size();
        }
        const tbs_plan *p = plans[s2];
        for (int l = 0; l < p->nline; ++l)
            lines.push_back({p->h_linesize[l], make_int4(s2, p->h_linepoint[l], p->h_linesize[l], p->h_woff[l])});
        for (int g = 0; g < p->ngroups; ++g) gtab.push_back(make_int2(s2, g));
    }
    std::stable_sort(lines.begin() + nlA, lines.end(), longest_first);
    const size_t ssz = b->prec == TBS_FP32 ? sizeof(SysDev<float>) : sizeof(SysDev<double>);
    const size_t nl = lines.size(), ng = gtab.size(), ns = (size_t)nplans;
    // grow the device tables (after the work that reads them has finished);
    // on any failure below the batch is left empty, so launches are no-ops
    b->plans.clear();
    b->nl = b->ng = 0;
    TBS_TRY(batch_grow(b, ns, nl, ng));
    // stage through pinned memory; the previous upload must have left it
    const size_t o_l = (ns * ssz + 15) / 16 * 16, o_g = o_l + nl * sizeof(int4);
    TBS_HIP(hipEventSynchronize(b->staged));
    for (size_t s2 = 0; s2 < ns; ++s2) {
        if (b->prec == TBS_FP32) { SysDev<float> d = make_sysdev<float>(plans[s2]); std::memcpy(b->h_stage + ssz * s2, &d, ssz); }
        else { SysDev<double> d = make_sysdev<double>(plans[s2]); std::memcpy(b->h_stage + ssz * s2, &d, ssz); }
    }
    for (size_t i = 0; i < nl; ++i) std::memcpy(b->h_stage + o_l + i * sizeof(int4), &lines[i].second, sizeof(int4));
    if (ng) std::memcpy(b->h_stage + o_g, gtab.data(), ng * sizeof(int2));
    if (ns) TBS_HIP(hipMemcpyAsync(b->d_sys, b->h_stage, ns * ssz, hipMemcpyHostToDevice, st));
    if (nl) TBS_HIP(hipMemcpyAsync(b->d_ltab, b->h_stage + o_l, nl * sizeof(int4), hipMemcpyHostToDevice, st));
    if (ng) TBS_HIP(hipMemcpyAsync(b->d_gtab, b->h_stage + o_g, ng * sizeof(int2), hipMemcpyHostToDevice, st));
    TBS_HIP(hipEventRecord(b->staged, st));
    b->plans.assign(plans, plans + nplans);
    b->nl = (int)nl;
    b->ng = (int)ng;
    b->nsA = nsA;
    b->nlA = (int)nlA;
    b->ngA = (int)ngA;
    return TBS_OK;
}

int tbs_batch_autotune(tbs_batch *b, int reps, int *chosen, double *us, int cap) {
    TBS_REQUIRE(b && reps >= 1, TBS_ERR_ARG, "bad arguments");
    TBS_REQUIRE(!b->plans.empty(), TBS_ERR_STATE, "batch has no systems");
    TBS_TRY(tbs::use_ctx(b->ctx));
    std::vector<std::pair<double *, long>> zv;
    const size_t tf = tf_size(b->prec);
    for (tbs_plan *p : b->plans) {
        const long nvec = (long)p->nle * p->nv;
        zv.push_back({p->d_U, nvec});
        zv.push_back({p->d_rhs, nvec});
        zv.push_back({(double *)p->d_g, (long)(nvec * tf / 8)});
    }
    const int ns = (int)b->plans.size();
    int tb = 0, rb = 0;
    TBS_TRY(autotune_kernels(b->ctx, b->nv, b->prec, b->d_sys, ns, b->d_ltab, b->nl, b->d_gtab, b->ng, zv, reps,
                             &tb, b->t_us, &rb, b->r_us));
    if (tb != b->variant || rb != b->rvariant) {
        for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        b->variant = tb;
        b->rvariant = rb;
    }
    // the two-stream split of eager sweeps, timed both ways with the chosen
    // kernels (a batch of one system has no halves)
    hipStream_t st = b->ctx->stream;
    b->pass_us[0] = b->pass_us[1] = 0.0;
    if (b->nsA > 0 && b->nsA < ns) {
        for (int sp2 = 0; sp2 < 2; ++sp2)
            TBS_TRY(time_calls(st, 1, [&]() { return batch_pass(b, st, reps, sp2 == 1); }, &b->pass_us[sp2]));
        b->pass_us[0] /= reps;
        b->pass_us[1] /= reps;
        b->split = b->pass_us[1] < b->pass_us[0];
    } else {
        b->split = false;
    }
    TBS_TRY(k_reset(b->prec, st, b->d_sys, ns));
    TBS_HIP(hipStreamSynchronize(st));
    if (chosen) *chosen = tb;
    for (int v = 0; us && v < cap && v < kTuneThomas; ++v) us[v] = b->t_us[v];
    return TBS_OK;
}

int tbs_batch_tuning(tbs_batch *b, tbs_tuning *out) {
    TBS_REQUIRE(b && out, TBS_ERR_ARG, "null argument");
    fill_tuning(b->nv, b->prec, b->variant, b->rvariant, b->split ? 1 : 0, b->t_us, b->r_us, b->pass_us, out);
    return TBS_OK;
}

int tbs_batch_set_tuning(tbs_batch *b, const tbs_tuning *t) {
    TBS_REQUIRE(b && t, TBS_ERR_ARG, "null argument");
    TBS_REQUIRE(t->thomas_variant >= 0 && t->thomas_variant < thomas_variants(b->nv, b->prec), TBS_ERR_ARG,
                "Thomas variant %d not built", t->thomas_variant);
    TBS_REQUIRE(t->residual_variant >= 0 && t->residual_variant < residual_variants(b->nv), TBS_ERR_ARG,
                "residual variant %d not built", t->residual_variant);
    if (t->thomas_variant != b->variant || t->residual_variant != b->rvariant) {
        for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
        b->graphs.clear();
        b->variant = t->thomas_variant;
        b->rvariant = t->residual_variant;
    }
    b->split = t->two_stream != 0;
    return TBS_OK;
}

int tbs_batch_launch_sweeps(tbs_batch *b, int nsweeps, int max_sweeps, double rtol) {
    TBS_REQUIRE(b && nsweeps >= 0 && max_sweeps >= 0 && max_sweeps < kHistCap, TBS_ERR_ARG, "bad arguments");
    TBS_TRY(tbs::use_ctx(b->ctx));
    if (b->plans.empty()) return TBS_OK;
    return batch_sweeps(b, b->ctx->stream, nsweeps, max_sweeps, rtol, b->split);
}

int tbs_batch_states(tbs_batch *b, int *states) {
    TBS_REQUIRE(b && (states || b->plans.empty()), TBS_ERR_ARG, "null argument");
    TBS_TRY(tbs::use_ctx(b->ctx));
    const int ns = (int)b->plans.size();
    if (ns == 0) return TBS_OK;
    hipStream_t st = b->ctx->stream;
    if (b->prec == TBS_FP32)
        hipLaunchKernelGGL((gather_states_kernel<float>), dim3((4 * ns + 255) / 256), dim3(256), 0, st,
                           (const SysDev<float> *)b->d_sys, ns, b->d_states);
    else
        hipLaunchKernelGGL((gather_states_kernel<double>), dim3((4 * ns + 255) / 256), dim3(256), 0, st,
                           (const SysDev<double> *)b->d_sys, ns, b->d_states);
    TBS_LAUNCHED("gather_states_kernel");
    TBS_HIP(hipMemcpyAsync(b->h_states, b->d_states, 4 * sizeof(int) * ns, hipMemcpyDeviceToHost, st));
    TBS_HIP(hipStreamSynchronize(st));
    std::memcpy(states, b->h_states, 4 * sizeof(int) * ns);
    return TBS_OK;
}

int tbs_batch_destroy(tbs_batch *b) {
    if (!b) return TBS_OK;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
    if (b->d_sys) hipFree(b->d_sys);
    if (b->d_ltab) hipFree(b->d_ltab);
    if (b->d_gtab) hipFree(b->d_gtab);
    if (b->d_states) hipFree(b->d_states);
    if (b->h_states) hipHostFree(b->h_states);
    if (b->h_stage) hipHostFree(b->h_stage);
    if (b->staged) hipEventDestroy(b->staged);
    delete b;
    return TBS_OK;
}

int tbs_batch_solve(tbs_batch *b, const double *const *rhs, double *const *U, int max_sweeps,
                    double rtol, int sweeps_per_check, double *hist, int *nsweeps) {
    TBS_REQUIRE(b && rhs && U, TBS_ERR_ARG, "null argument");
    TBS_REQUIRE(!b->plans.empty(), TBS_ERR_STATE, "batch has no systems");
    TBS_REQUIRE(max_sweeps >= 0 && max_sweeps < kHistCap, TBS_ERR_ARG, "bad max_sweeps");
    TBS_TRY(tbs::use_ctx(b->ctx));
    const int ns = (int)b->plans.size();
    hipStream_t st = b->ctx->stream;
    if (sweeps_per_check < 1) sweeps_per_check = 1;
    for (int s = 0; s < ns; ++s) TBS_REQUIRE(rhs[s] && U[s], TBS_ERR_ARG, "null vector for system %d", s);
    TBS_TRY(k_reset(b->prec, st, b->d_sys, ns));
    for (int s = 0; s < ns; ++s) {
        tbs_plan *p = b->plans[s];
        const long total = (long)p->nle * p->nv;
        hipLaunchKernelGGL(gather_vec_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                           p->d_lines, rhs[s], p->d_rhs);
        hipLaunchKernelGGL(gather_u0_kernel, dim3(gs_grid(total)), dim3(256), 0, st, total, p->nv,
                           p->d_lines, (const double *)U[s], p->d_U, p->d_state + 3);
        TBS_LAUNCHED("batch gather");
    }
    // J U skipped for the systems whose initial U is all zeros
    TBS_TRY(k_residual(b->nv, b->prec, st, b->d_sys, b->d_gtab, b->ng, 0, 1 | 4, nullptr, nullptr, b->rvariant));
    TBS_TRY(k_final(b->prec, st, b->d_sys, ns, max_sweeps, rtol, 1));
    auto key = std::make_pair(sweeps_per_check, std::make_pair(max_sweeps, rtol));
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        hipGraphExec_t ge = nullptr;
        auto all = [b, st, sweeps_per_check, max_sweeps, rtol]() {
            return batch_sweeps(b, st, sweeps_per_check, max_sweeps, rtol, false);
        };
        TBS_TRY(capture_sweeps(st, 1, &ge, all));
        it = b->graphs.emplace(key, ge).first;
    }
    int done = 0;
    std::vector<int> st_h(4 * ns);
    while (done < max_sweeps) {
        TBS_HIP(hipGraphLaunch(it->second, st));
        done += sweeps_per_check;
        bool any = false;
        TBS_TRY(tbs_batch_states(b, st_h.data()));
        for (int s = 0; s < ns; ++s) any |= st_h[4 * s + 1] != 0;
        if (!any) break;
    }
    for (int s = 0; s < ns; ++s) {
        TBS_TRY(tbs_plan_end(b->plans[s], U[s], nullptr));
        int count = 0;
        TBS_TRY(tbs_plan_history(b->plans[s], hist ? hist + (size_t)s * (max_sweeps + 1) : nullptr,
                                 max_sweeps + 1, &count));
        if (nsweeps) nsweeps[s] = count > 0 ? count - 1 : 0;
    }
    return TBS_OK;
}

}  // extern "C"
