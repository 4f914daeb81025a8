// server.hip -- continuous-batching LIJ solve service on one GPU.
//
// Requests (seeded structured systems, FP32-factor or FP64 plans) wait in a
// FIFO queue.  Each engine step
//   1. enqueues K sweeps of the current batch on the main stream,
//   2. admits queued requests whose HBM footprint fits the budget: plan
//      tables, device generation of the Jacobian, factorisation and the
//      initial residual run on a second (admission) stream, overlapping the
//      batch's sweeps on the device,
//   3. waits for the batch's sweeps and reads every system's state,
//   4. retires finished systems (U copied to host, HBM returned), and
//   5. moves admitted requests whose admission work has completed into the
//      batch, rebuilding only the descriptor tables.
// A request's arithmetic depends on its own data alone (per-line Thomas,
// per-element residual, per-system fixed-order norm), so its result is
// bitwise the same whatever it shares the batch with.
#include "tbs_common.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;

struct Request {
    long id = 0;
    tbs_server_request spec{};
    int nelem = 0, nintface = 0;
    size_t footprint = 0;
    // mesh (host; the plan keeps its own device copies)
    std::vector<int> epoint, ef, fc, linesize, linepoint, lines, lineface;
    tbs_plan *plan = nullptr;
    double *d_rhs = nullptr, *d_U = nullptr;
    hipEvent_t ready = nullptr;
    // results
    std::vector<double> U, hist;
    int nsweeps = 0, status = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bool done = false;
};

size_t request_footprint(const tbs_mesh_host &mh, int prec, int nvar) {
    size_t plan = 0;
    if (tbs_plan_footprint(&mh, prec, &plan) != TBS_OK) return (size_t)-1;
    const size_t bs = (size_t)nvar * nvar * 8;
    const size_t jac = bs * ((size_t)mh.nelem + 2 * (size_t)mh.nintface);   // generated D, O1, O2
    const size_t work = 3 * bs * (size_t)mh.nelem;                          // factor workspace
    const size_t vec = 2 * (size_t)nvar * mh.nelem * 8;                     // rhs, U
    return plan + jac + work + vec;
}

// Structured NL x NK mesh of a request, in the request's host arrays.
int build_mesh(Request *r) {
    const int NL = r->spec.NL, NK = r->spec.NK;
    int nif = 0, n = 0, eft = 0, nl = 0, nle = 0;
    TBS_TRY(tbs_structured_sizes(NL, NK, &n, &nif, &eft, &nl, &nle));
    r->nelem = n;
    r->nintface = nif;
    r->epoint.resize(n + 1);
    r->ef.resize(eft);
    r->fc.resize(2 * (size_t)nif);
    r->linesize.resize(nl);
    r->linepoint.resize(nl + 1);
    r->lines.resize(nle);
    r->lineface.resize(nle);
    return tbs_structured_mesh(NL, NK, r->epoint.data(), r->ef.data(), r->fc.data(), r->linesize.data(),
                               r->linepoint.data(), r->lines.data(), r->lineface.data());
}

tbs_mesh_host mesh_view(const Request &r, int nvar) {
    tbs_mesh_host mh{};
    mh.nvar = nvar;
    mh.nelem = r.nelem;
    mh.nintface = r.nintface;
    mh.eftot = (int)r.ef.size();
    mh.nline = (int)r.linesize.size();
    mh.nlineelem = (int)r.lines.size();
    mh.epoint = r.epoint.data();
    mh.ef = r.ef.data();
    mh.fc = r.fc.data();
    mh.linesize = r.linesize.data();
    mh.linepoint = r.linepoint.data();
    mh.lines = r.lines.data();
    mh.lineface = r.lineface.data();
    return mh;
}

}  // namespace

struct tbs_server {
    tbs_ctx *ctx = nullptr;       // batch stream
    tbs_ctx *adm = nullptr;       // admission stream (same device)
    size_t budget = 0, used = 0;
    int prec = TBS_FP32, max_sweeps = 100, k = 4;
    double rtol = 1e-10;
    long next_id = 1;
    clk::time_point t0;
    std::deque<std::unique_ptr<Request>> queue;
    std::vector<std::unique_ptr<Request>> admitting, active;
    std::map<long, std::unique_ptr<Request>> finished;
    std::vector<long> completed_unpolled;
    tbs_batch *batch = nullptr;
    bool dirty = false;
    double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
};

namespace {

void free_request_device(Request *r) {
    if (r->plan) tbs_plan_destroy(r->plan);
    r->plan = nullptr;
    if (r->d_rhs) hipFree(r->d_rhs);
    if (r->d_U) hipFree(r->d_U);
    r->d_rhs = r->d_U = nullptr;
    if (r->ready) hipEventDestroy(r->ready);
    r->ready = nullptr;
}

int admit(tbs_server *s, Request *r) {
    const tbs_mesh_host mh = mesh_view(*r, s->ctx->nvar);
    const int nv = mh.nvar;
    const size_t bs = (size_t)nv * nv;
    int rc = tbs_plan_create(s->adm, &mh, s->prec, &r->plan);
    if (rc) return rc;
    double *D = nullptr, *O1 = nullptr, *O2 = nullptr;
    const size_t nD = bs * r->nelem, nO = bs * (size_t)std::max(r->nintface, 1);
    hipError_t e = hipMalloc((void **)&D, nD * 8);
    if (!e) e = hipMalloc((void **)&O1, nO * 8);
    if (!e) e = hipMalloc((void **)&O2, nO * 8);
    if (!e) e = hipMalloc((void **)&r->d_rhs, (size_t)nv * r->nelem * 8);
    if (!e) e = hipMalloc((void **)&r->d_U, (size_t)nv * r->nelem * 8);
    if (e) {
        hipGetLastError();
        if (D) hipFree(D);
        if (O1) hipFree(O1);
        if (O2) hipFree(O2);
        tbs::set_error("admission allocation refused");
        return TBS_ERR_NOMEM;
    }
    do {
        if ((rc = tbs_generate_device(s->adm, r->spec.seed, nv, r->nelem, r->nintface, r->spec.shift, D, O1,
                                      O2, r->d_rhs)))
            break;
        if ((rc = tbs_memset(s->adm, r->d_U, 0, (size_t)nv * r->nelem * 8))) break;
        if ((rc = tbs_plan_factor(r->plan, D, O1, O2))) break;      // synchronises the admission stream
        if ((rc = tbs_plan_begin(r->plan, r->d_rhs, r->d_U))) break;
        if (hipEventCreateWithFlags(&r->ready, hipEventDisableTiming) ||
            hipEventRecord(r->ready, (hipStream_t)tbs_ctx_stream(s->adm))) {
            rc = TBS_ERR_STATE;
            break;
        }
    } while (0);
    hipFree(D);
    hipFree(O1);
    hipFree(O2);
    return rc;
}

int rebuild_batch(tbs_server *s) {
    if (s->batch) {
        tbs_batch_destroy(s->batch);
        s->batch = nullptr;
    }
    if (s->active.empty()) return TBS_OK;
    std::vector<tbs_plan *> plans;
    for (auto &r : s->active) plans.push_back(r->plan);
    return tbs_batch_create(s->ctx, plans.data(), (int)plans.size(), &s->batch);
}

}  // namespace

extern "C" {

int tbs_server_create(tbs_ctx *ctx, size_t hbm_budget, int precision, double rtol, int max_sweeps,
                      int sweeps_per_step, tbs_server **out) {
    TBS_TRY(tbs::use_ctx(ctx));
    TBS_REQUIRE(out && (precision == TBS_FP32 || precision == TBS_FP64) && rtol >= 0 && max_sweeps >= 1 &&
                    max_sweeps < 4096 && sweeps_per_step >= 1, TBS_ERR_ARG, "bad server arguments");
    *out = nullptr;
    std::unique_ptr<tbs_server> s(new tbs_server());
    s->ctx = ctx;
    TBS_TRY(tbs_ctx_create(ctx->device, ctx->nvar, &s->adm));
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        TBS_HIP(hipMemGetInfo(&fr, &tot));
        hbm_budget = fr / 10 * 8;
    }
    s->budget = hbm_budget;
    s->prec = precision;
    s->rtol = rtol;
    s->max_sweeps = max_sweeps;
    s->k = sweeps_per_step;
    s->t0 = clk::now();
    *out = s.release();
    return TBS_OK;
}

int tbs_server_destroy(tbs_server *s) {
    if (!s) return TBS_OK;
    hipSetDevice(s->ctx->device);
    hipDeviceSynchronize();
    if (s->batch) tbs_batch_destroy(s->batch);
    for (auto &r : s->active) free_request_device(r.get());
    for (auto &r : s->admitting) free_request_device(r.get());
    for (auto &kv : s->finished) free_request_device(kv.second.get());
    tbs_ctx_destroy(s->adm);
    delete s;
    return TBS_OK;
}

int tbs_server_footprint(tbs_server *s, const tbs_server_request *req, size_t *bytes) {
    TBS_REQUIRE(s && req && bytes, TBS_ERR_ARG, "null argument");
    Request r;
    r.spec = *req;
    TBS_TRY(build_mesh(&r));
    *bytes = request_footprint(mesh_view(r, s->ctx->nvar), s->prec, s->ctx->nvar);
    return TBS_OK;
}

int tbs_server_submit(tbs_server *s, const tbs_server_request *req, long *id) {
    TBS_REQUIRE(s && req && id, TBS_ERR_ARG, "null argument");
    std::unique_ptr<Request> r(new Request());
    r->spec = *req;
    TBS_TRY(build_mesh(r.get()));
    r->footprint = request_footprint(mesh_view(*r, s->ctx->nvar), s->prec, s->ctx->nvar);
    TBS_REQUIRE(r->footprint <= s->budget, TBS_ERR_NOMEM, "request needs %zu bytes, budget is %zu",
                r->footprint, s->budget);
    r->id = s->next_id++;
    r->t_submit = s->now();
    *id = r->id;
    s->queue.push_back(std::move(r));
    return TBS_OK;
}

int tbs_server_step(tbs_server *s, int *nactive, int *nqueued, int *nadmitting) {
    TBS_REQUIRE(s, TBS_ERR_ARG, "null server");
    TBS_TRY(tbs::use_ctx(s->ctx));
    // 1. sweeps of the current batch (asynchronous)
    if (s->batch) TBS_TRY(tbs_batch_launch_sweeps(s->batch, s->k, s->max_sweeps, s->rtol));
    // 2. admission, FIFO with skip-ahead for requests that fit
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Request *r = it->get();
        if (s->used + r->footprint > s->budget) { ++it; continue; }
        s->used += r->footprint;
        r->t_admit = s->now();
        int rc = admit(s, r);
        if (rc) {
            s->used -= r->footprint;
            free_request_device(r);
            r->status = -1;
            r->done = true;
            r->t_done = s->now();
            s->completed_unpolled.push_back(r->id);
            s->finished[r->id] = std::move(*it);
            it = s->queue.erase(it);
            continue;
        }
        s->admitting.push_back(std::move(*it));
        it = s->queue.erase(it);
    }
    // 3. states of the batch
    if (s->batch) {
        std::vector<int> st(4 * s->active.size());
        TBS_TRY(tbs_batch_states(s->batch, st.data()));
        // 4. retire
        std::vector<std::unique_ptr<Request>> keep;
        for (size_t i = 0; i < s->active.size(); ++i) {
            std::unique_ptr<Request> &r = s->active[i];
            if (st[4 * i + 1]) { keep.push_back(std::move(r)); continue; }
            const int nv = s->ctx->nvar;
            r->status = st[4 * i + 2];
            r->U.resize((size_t)nv * r->nelem);
            r->hist.resize(s->max_sweeps + 1);
            int cnt = 0;
            TBS_TRY(tbs_plan_end(r->plan, r->d_U, nullptr));
            TBS_TRY(tbs_plan_history(r->plan, r->hist.data(), (int)r->hist.size(), &cnt));
            TBS_TRY(tbs_memcpy_d2h(s->adm, r->U.data(), r->d_U, r->U.size() * 8));
            r->hist.resize(cnt);
            r->nsweeps = cnt > 0 ? cnt - 1 : 0;
            r->t_done = s->now();
            r->done = true;
            free_request_device(r.get());
            s->used -= r->footprint;
            s->completed_unpolled.push_back(r->id);
            s->finished[r->id] = std::move(r);
            s->dirty = true;
        }
        s->active.swap(keep);
    }
    // 5. join admitted requests whose admission work is complete
    for (auto it = s->admitting.begin(); it != s->admitting.end();) {
        if (hipEventQuery((*it)->ready) == hipSuccess) {
            s->active.push_back(std::move(*it));
            it = s->admitting.erase(it);
            s->dirty = true;
        } else {
            ++it;
        }
    }
    hipGetLastError();   // hipEventQuery reports "not ready" as an error code
    if (s->dirty) {
        TBS_TRY(rebuild_batch(s));
        s->dirty = false;
    }
    if (nactive) *nactive = (int)s->active.size();
    if (nqueued) *nqueued = (int)s->queue.size();
    if (nadmitting) *nadmitting = (int)s->admitting.size();
    return TBS_OK;
}

int tbs_server_poll(tbs_server *s, long *ids, int cap, int *n) {
    TBS_REQUIRE(s && n, TBS_ERR_ARG, "null argument");
    int k = 0;
    std::vector<long> rest;
    for (long id : s->completed_unpolled) {
        if (ids && k < cap) ids[k++] = id;
        else rest.push_back(id);
    }
    s->completed_unpolled.swap(rest);
    *n = k;
    return TBS_OK;
}

int tbs_server_result(tbs_server *s, long id, double *U, size_t ucap, double *hist, int hcap,
                      int *nsweeps, int *status, double *times) {
    TBS_REQUIRE(s, TBS_ERR_ARG, "null server");
    auto it = s->finished.find(id);
    TBS_REQUIRE(it != s->finished.end(), TBS_ERR_ARG, "request %ld is not finished", id);
    Request *r = it->second.get();
    if (U) {
        TBS_REQUIRE(ucap >= r->U.size(), TBS_ERR_ARG, "U capacity %zu < %zu", ucap, r->U.size());
        std::memcpy(U, r->U.data(), r->U.size() * 8);
    }
    if (hist) std::memcpy(hist, r->hist.data(), std::min<size_t>(hcap, r->hist.size()) * 8);
    if (nsweeps) *nsweeps = r->nsweeps;
    if (status) *status = r->status;
    if (times) {
        times[0] = r->t_submit;
        times[1] = r->t_admit;
        times[2] = r->t_done;
    }
    return TBS_OK;
}

int tbs_server_release(tbs_server *s, long id) {
    TBS_REQUIRE(s, TBS_ERR_ARG, "null server");
    s->finished.erase(id);
    return TBS_OK;
}

int tbs_server_info(tbs_server *s, size_t *budget, size_t *used, double *now) {
    TBS_REQUIRE(s, TBS_ERR_ARG, "null server");
    if (budget) *budget = s->budget;
    if (used) *used = s->used;
    if (now) *now = s->now();
    return TBS_OK;
}

}  // extern "C"
