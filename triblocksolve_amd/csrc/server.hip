// server.hip -- continuous-batching LIJ solve service on one GPU.
//
// Requests (a caller's system -- mesh, Jacobian blocks and rhs in the
// reference's layout, from host or device memory -- or a seeded structured
// system the server generates; FP32-factor or FP64 plans) wait in a FIFO
// queue.  The server owns an HBM arena, sized to the budget when the server
// is created (grown in segments of at least kSegment bytes only if that
// first allocation was refused); every admitted request gets one slab of it
// (first fit), from which its plan, its rhs/U and its admission temporaries
// are carved, so serving makes no hipMalloc/hipFree calls.  Each engine step
//   1. reads the states of the running batch (waits for its last sweeps),
//   2. drops finished systems from the batch and adds admitted ones whose
//      set-up has completed (descriptor tables rewritten in place),
//   3. enqueues K sweeps of the batch,
//   4. while those run, retires the finished systems on a third stream (U
//      scattered to element order and copied to the host, slab returned), and
//   5. admits queued requests that fit: each gets its slab and goes to the
//      admission thread, which builds its plan tables and enqueues the
//      Jacobian (copied from the caller's arrays or generated), the
//      factorisation and the initial residual on the admission stream, so
//      neither that host work nor the device work holds up the steps.
// Everything a step could otherwise allocate (pinned result buffers, table
// staging, the batch's tables) is allocated when the server is created.
// A request's arithmetic depends on its own data alone (per-line Thomas,
// per-element residual, per-system fixed-order norm), so its result is
// bitwise the same whatever it shares the batch with.
#include "tbs_common.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;

constexpr size_t kAlign = tbs::kRegionAlign;
constexpr size_t kSlack = 16 * kAlign;   // bump-alignment reserve per request
// a request behind the queue head may be admitted first while the head has
// been passed over fewer than this many times
constexpr int kMaxSkips = 16;
constexpr size_t kSegment = (size_t)16 << 30;   // arena growth step

size_t align_up(size_t b) { return (std::max<size_t>(b, 1) + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over the ranges released into it, coalescing
// adjacent free ranges.
class Arena {
  public:
    bool alloc(size_t n, size_t *off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const size_t rest = it->second - n, at = it->first + n;
            free_.erase(it);
            if (rest) free_[at] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        if (!n) return;
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && off + n == nx->first) {
            n += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {
                pv->second += n;
                return;
            }
        }
        free_[off] = n;
    }

  private:
    std::map<size_t, size_t> free_;   // offset -> length
};

struct Request {
    long id = 0;
    tbs_server_request spec{};
    // a caller's system: its Jacobian and rhs, host copies taken at submit
    // (where == TBS_HOST) or the caller's device arrays, read at admission
    bool caller = false;
    int where = TBS_HOST;
    std::vector<double> hD, hO1, hO2, hrhs;
    const double *uD = nullptr, *uO1 = nullptr, *uO2 = nullptr, *urhs = nullptr;
    int nelem = 0, nintface = 0, eftot = 0, nlineelem = 0;
    size_t footprint = 0;   // slab bytes while admitting
    size_t persist = 0;     // slab bytes kept once admitted (plan, rhs, U)
    size_t slab_off = 0, slab_len = 0;
    char *slab_ptr = nullptr;   // device address of the slab (set before the hand-off)
    int skipped = 0;
    // admission thread -> serving thread: 0 set-up pending, 1 set-up enqueued
    // (ready recorded), -1 set-up failed (slab no longer in use; err holds
    // the admission thread's error message)
    std::atomic<int> setup{0};
    std::string err;
    std::vector<int> epoint, ef, fc, linesize, linepoint, lines, lineface;
    tbs_plan *plan = nullptr;
    double *d_rhs = nullptr, *d_U = nullptr;   // element order, in the slab
    double *tD = nullptr, *tO1 = nullptr, *tO2 = nullptr;   // generated Jacobian (slab temporaries)
    hipEvent_t ready = nullptr;
    double *hU = nullptr;           // result U (pinned, from the server's pool)
    size_t nU = 0, hU_cls = 0;
    double *hH = nullptr;           // residual history (pinned, from the pool)
    size_t nH = 0, hH_cls = 0;
    int nsweeps = 0, status = 0;
    // timeline (seconds on the server's clock): submitted, admitted (set-up
    // enqueued), set-up seen complete (joins the batch), convergence seen by
    // a step, result copies enqueued, copies seen complete (result ready)
    double t_submit = 0, t_admit = 0, t_ready = 0, t_conv = 0, t_retire = 0, t_done = 0;
};

// A seeded request's grid: its sizes and line arrays only; the plan
// generates the element and face tables on the device.
int build_mesh(Request *r) {
    const int NL = r->spec.NL, NK = r->spec.NK;
    int nif = 0, n = 0, eft = 0, nl = 0, nle = 0;
    TBS_TRY(tbs_structured_sizes(NL, NK, &n, &nif, &eft, &nl, &nle));
    r->nelem = n;
    r->nintface = nif;
    r->eftot = eft;
    r->nlineelem = nle;
    r->linesize.assign(nl, NK);
    r->linepoint.resize(nl + 1);
    for (int l = 0; l <= nl; ++l) r->linepoint[l] = l * NK;
    return TBS_OK;
}

tbs_mesh_host mesh_view(const Request &r, int nvar) {
    tbs_mesh_host mh{};
    mh.nvar = nvar;
    mh.nelem = r.nelem;
    mh.nintface = r.nintface;
    mh.eftot = r.eftot;
    mh.nline = (int)r.linesize.size();
    mh.nlineelem = r.nlineelem;
    mh.linesize = r.linesize.data();
    mh.linepoint = r.linepoint.data();
    if (r.caller) {   // a seeded request's grid has no host element/face tables
        mh.epoint = r.epoint.data();
        mh.ef = r.ef.data();
        mh.fc = r.fc.data();
        mh.lines = r.lines.data();
        mh.lineface = r.lineface.data();
    }
    return mh;
}

// Slab sizes: persistent part (plan arrays, rhs, U) and the whole slab
// (plus generated D/O1/O2 and the factorisation workspace).
int request_sizes(const Request &r, int prec, int nvar, size_t *persist, size_t *total) {
    const tbs_mesh_host mh = mesh_view(r, nvar);
    size_t plan = 0;
    TBS_TRY(tbs::plan_footprint(&mh, prec, &plan, false, !r.caller));   // checked at submit
    const size_t bs = (size_t)nvar * nvar * 8, vec = (size_t)nvar * r.nelem * 8;
    *persist = align_up(plan + 2 * align_up(vec) + kSlack);
    *total = *persist + align_up(bs * r.nelem) + 2 * align_up(bs * std::max(r.nintface, 1)) +
             tbs::plan_factor_workspace(nvar, r.nelem) + kSlack;
    return TBS_OK;
}

enum { T_WAIT, T_MEMBERS, T_LAUNCH, T_RETIRE, T_ADMIT, T_PLAN, T_SETUP, T_NPHASE };
enum { C_STEPS, C_ADMITTED, C_RETIRED, C_REBUILDS, C_NCOUNT };

}  // namespace

// Pinned host buffers for results, in power-of-two size classes, reused
// from request to request (a result copied into pageable memory costs page
// faults and a staged copy on the serving thread).
struct PinnedPool {
    std::map<size_t, std::vector<double *>> free;
    std::vector<double *> all;
    int get(size_t bytes, double **p, size_t *cls) {
        size_t c = 4096;
        while (c < bytes) c <<= 1;
        std::vector<double *> &v = free[c];
        if (!v.empty()) {
            *p = v.back();
            v.pop_back();
        } else {
            TBS_HIP(hipHostMalloc((void **)p, c, hipHostMallocDefault));
            all.push_back(*p);
        }
        *cls = c;
        return TBS_OK;
    }
    void put(double *p, size_t cls) {
        if (p) free[cls].push_back(p);
    }
    // n buffers of every class from 4 KiB up to max_bytes, allocated now
    int prefill(size_t max_bytes, int n) {
        for (size_t c = 4096; c <= max_bytes; c <<= 1)
            for (int i = 0; i < n; ++i) {
                double *p = nullptr;
                TBS_HIP(hipHostMalloc((void **)&p, c, hipHostMallocDefault));
                all.push_back(p);
                free[c].push_back(p);
            }
        return TBS_OK;
    }
    void clear() {
        for (double *p : all) hipHostFree(p);
        all.clear();
        free.clear();
    }
};

struct tbs_server {
    tbs_ctx *ctx = nullptr;   // batch stream
    tbs_ctx *adm = nullptr;   // admission stream (same device)
    tbs_ctx *ret = nullptr;   // retirement stream (same device)
    size_t budget = 0, used = 0;
    int prec = TBS_FP32, max_sweeps = 100, k = 4;
    double rtol = 1e-10;
    long next_id = 1;
    clk::time_point t0;
    // arena segments; slab offsets are in a virtual space where segment i
    // starts at seg_off[i] (segments never coalesce with each other)
    std::vector<char *> seg_base;
    std::vector<size_t> seg_off, seg_len;
    size_t arena_bytes = 0;
    Arena arena;
    char *addr(size_t off) const {
        for (size_t i = 0; i < seg_base.size(); ++i)
            if (off >= seg_off[i] && off < seg_off[i] + seg_len[i]) return seg_base[i] + (off - seg_off[i]);
        return nullptr;
    }
    std::deque<std::unique_ptr<Request>> queue;
    std::vector<std::unique_ptr<Request>> admitting, active;
    std::vector<std::unique_ptr<Request>> retiring;   // results being copied out on the retirement stream
    std::map<long, std::unique_ptr<Request>> finished;
    std::vector<long> completed_unpolled;
    tbs_batch *batch = nullptr;
    PinnedPool pinned;
    double t_phase[T_NPHASE] = {};
    long count[C_NCOUNT] = {};
    double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
    // admission thread: requests handed over by the steps (owned by
    // `admitting`), the thread's own host-time counters (nanoseconds)
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request *> handoff;
    bool stop = false;
    std::atomic<long long> ns_plan{0}, ns_setup{0};
};

namespace {

void drop_device(tbs_server *s, Request *r) {
    if (r->plan) tbs_plan_destroy(r->plan);   // region-backed: frees no device memory
    r->plan = nullptr;
    if (r->ready) hipEventDestroy(r->ready);
    r->ready = nullptr;
    if (r->slab_len) s->arena.release(r->slab_off, r->slab_len);
    r->slab_len = 0;
    r->d_rhs = r->d_U = nullptr;
}

// Enqueue a request's set-up on the admission stream, inside its slab:
// plan tables, device generation of its Jacobian and rhs, U = 0.  Its
// factorisation and first residual follow in admit_finish, together with the
// other requests admitted in the same step.
int admit_setup(tbs_server *s, Request *r) {
    const tbs_mesh_host mh = mesh_view(*r, s->ctx->nvar);
    const int nv = mh.nvar;
    const size_t bs = (size_t)nv * nv;
    tbs_ctx *a = s->adm;
    a->region = r->slab_ptr;
    a->region_size = r->slab_len;
    a->region_used = 0;
    int rc = TBS_OK;
    do {
        const auto tp0 = clk::now();
        // (a caller's mesh was checked at submit; a seeded grid's tables are
        // generated on the device)
        rc = tbs::plan_create(a, &mh, s->prec, &r->plan, false, r->caller ? 0 : r->spec.NK);
        s->ns_plan += std::chrono::duration_cast<std::chrono::nanoseconds>(clk::now() - tp0).count();
        if (rc) break;
        if ((rc = tbs::dev_alloc(a, (size_t)nv * r->nelem * 8, (void **)&r->d_rhs, nullptr))) break;
        if ((rc = tbs::dev_alloc(a, (size_t)nv * r->nelem * 8, (void **)&r->d_U, nullptr))) break;
        if (a->region_used > r->persist) { rc = TBS_ERR_STATE; tbs::set_error("slab layout overflow"); break; }
        a->region_used = r->persist;   // temporaries start at the persistent boundary
        if ((rc = tbs::dev_alloc(a, bs * r->nelem * 8, (void **)&r->tD, nullptr))) break;
        if ((rc = tbs::dev_alloc(a, bs * std::max(r->nintface, 1) * 8, (void **)&r->tO1, nullptr))) break;
        if ((rc = tbs::dev_alloc(a, bs * std::max(r->nintface, 1) * 8, (void **)&r->tO2, nullptr))) break;
        if (r->caller) {
            // the caller's Jacobian and rhs into the slab (rhs stays, in
            // element order, for the plan's gather)
            const bool dv = r->where == TBS_DEVICE;
            const hipMemcpyKind kind = dv ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
            const size_t nD = bs * r->nelem * 8, nO = bs * (size_t)r->nintface * 8, nr = (size_t)nv * r->nelem * 8;
            hipError_t e = hipMemcpyAsync(r->tD, dv ? r->uD : r->hD.data(), nD, kind, a->stream);
            if (!e && nO) e = hipMemcpyAsync(r->tO1, dv ? r->uO1 : r->hO1.data(), nO, kind, a->stream);
            if (!e && nO) e = hipMemcpyAsync(r->tO2, dv ? r->uO2 : r->hO2.data(), nO, kind, a->stream);
            if (!e) e = hipMemcpyAsync(r->d_rhs, dv ? r->urhs : r->hrhs.data(), nr, kind, a->stream);
            if (e) { rc = tbs::hip_fail(e, "request copy-in", __FILE__, __LINE__); break; }
        } else if ((rc = tbs_generate_device(a, r->spec.seed, nv, r->nelem, r->nintface, r->spec.shift, r->tD,
                                             r->tO1, r->tO2, r->d_rhs))) {
            break;
        }
        if ((rc = tbs_memset(a, r->d_U, 0, (size_t)nv * r->nelem * 8))) break;
    } while (0);
    a->region = nullptr;
    a->region_size = a->region_used = 0;
    if (rc) hipStreamSynchronize(a->stream);   // nothing of the slab may still be in use
    return rc;
}

// Factorise the requests set up in this step in one launch over all their
// lines (a small request's factorisation is a latency-bound chain per line
// on a few CUs; one launch runs them side by side), then their first
// residuals and set-up events.
int admit_finish(tbs_server *s, const std::vector<Request *> &rs) {
    tbs_ctx *a = s->adm;
    std::vector<tbs_plan *> plans;
    std::vector<const double *> D, O1, O2;
    for (Request *r : rs) {
        plans.push_back(r->plan);
        D.push_back(r->tD);
        O1.push_back(r->tO1);
        O2.push_back(r->tO2);
    }
    int rc = tbs_plans_factor(a, plans.data(), (int)plans.size(), D.data(), O1.data(), O2.data());
    for (size_t i = 0; i < rs.size() && rc == TBS_OK; ++i) {
        Request *r = rs[i];
        if ((rc = tbs_plan_begin(r->plan, r->d_rhs, r->d_U))) break;
        if (hipEventCreateWithFlags(&r->ready, hipEventDisableTiming) || hipEventRecord(r->ready, a->stream)) {
            rc = TBS_ERR_STATE;
            tbs::set_error("admission event");
        }
    }
    if (rc) hipStreamSynchronize(a->stream);   // nothing of the slabs may still be in use
    return rc;
}

// The admission thread: takes the requests the steps hand over, all those
// waiting at once, sets each up on the admission stream and factorises them
// in one launch, then publishes the outcome through Request::setup.
void admission_loop(tbs_server *s) {
    const bool usable = tbs::use_ctx(s->adm) == TBS_OK;   // (if not, every set-up fails)
    for (;;) {
        std::vector<Request *> got;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [s] { return s->stop || !s->handoff.empty(); });
            if (s->handoff.empty()) return;   // stop, nothing left
            got.assign(s->handoff.begin(), s->handoff.end());
            s->handoff.clear();
        }
        const auto t0 = clk::now();
        std::vector<Request *> ok;
        for (Request *r : got) {
            if (usable && admit_setup(s, r) == TBS_OK) {
                ok.push_back(r);
            } else {
                r->err = usable ? tbs_last_error() : "admission device could not be selected";
                r->setup.store(-1, std::memory_order_release);
            }
        }
        const int rc = ok.empty() ? TBS_OK : admit_finish(s, ok);
        s->ns_setup += std::chrono::duration_cast<std::chrono::nanoseconds>(clk::now() - t0).count();
        for (Request *r : ok) {
            if (rc != TBS_OK) r->err = tbs_last_error();
            r->setup.store(rc == TBS_OK ? 1 : -1, std::memory_order_release);
        }
    }
}

// Add a segment able to hold `need` bytes, within the budget.
int grow_arena(tbs_server *s, size_t need) {
    const size_t room = s->budget - std::min(s->budget, s->arena_bytes);
    size_t len = std::min(std::max(need, kSegment), room);
    TBS_REQUIRE(len >= need, TBS_ERR_NOMEM, "arena cannot grow by %zu bytes", need);
    char *base = nullptr;
    hipError_t e = hipMalloc((void **)&base, len);
    if (e) {
        hipGetLastError();
        tbs::set_error("server arena segment of %zu bytes refused", len);
        return TBS_ERR_NOMEM;
    }
    // a gap of one alignment unit keeps segments from coalescing in the free list
    const size_t off = s->seg_off.empty() ? 0 : s->seg_off.back() + s->seg_len.back() + kAlign;
    s->seg_base.push_back(base);
    s->seg_off.push_back(off);
    s->seg_len.push_back(len);
    s->arena_bytes += len;
    s->arena.release(off, len);
    return TBS_OK;
}

void finish(tbs_server *s, std::unique_ptr<Request> r) {
    r->t_done = s->now();
    s->completed_unpolled.push_back(r->id);
    const long id = r->id;
    s->finished[id] = std::move(r);
}

// Finish the retiring requests whose copies have completed (all of them,
// waiting, when wait is set); their slabs go back to the arena only then.
int reap(tbs_server *s, bool wait) {
    std::vector<std::unique_ptr<Request>> keep;
    int rc = TBS_OK;
    for (auto &r : s->retiring) {
        if (wait) {
            const hipError_t e = hipEventSynchronize(r->ready);
            if (e != hipSuccess && rc == TBS_OK) {
                tbs::set_error("retirement copy: %s", hipGetErrorString(e));
                rc = TBS_ERR_STATE;
            }
        } else if (hipEventQuery(r->ready) != hipSuccess) {
            keep.push_back(std::move(r));
            continue;
        }
        s->used -= r->slab_len;
        drop_device(s, r.get());
        s->count[C_RETIRED]++;
        finish(s, std::move(r));
    }
    hipGetLastError();   // hipEventQuery reports "not ready" as an error code
    s->retiring.swap(keep);
    return rc;
}

// A caller's system as a request: mesh checked as tbs_plan_create checks it,
// mesh arrays copied, Jacobian and rhs copied now (host) or at admission
// (device).
int caller_request(tbs_server *s, const tbs_mesh_host *m, std::unique_ptr<Request> *out) {
    TBS_REQUIRE(m, TBS_ERR_ARG, "null mesh");
    TBS_REQUIRE(m->nvar == s->ctx->nvar, TBS_ERR_ARG, "mesh nvar %d != server nvar %d", m->nvar, s->ctx->nvar);
    size_t plan_bytes = 0;
    TBS_TRY(tbs::plan_footprint(m, s->prec, &plan_bytes, true));
    std::unique_ptr<Request> r(new Request());
    r->caller = true;
    r->nelem = m->nelem;
    r->nintface = m->nintface;
    r->eftot = m->eftot;
    r->nlineelem = m->nlineelem;
    r->epoint.assign(m->epoint, m->epoint + m->nelem + 1);
    r->ef.assign(m->ef, m->ef + m->eftot);
    r->fc.assign(m->fc, m->fc + 2 * (size_t)m->nintface);
    r->linesize.assign(m->linesize, m->linesize + m->nline);
    r->linepoint.assign(m->linepoint, m->linepoint + m->nline + 1);
    r->lines.assign(m->lines, m->lines + m->nlineelem);
    r->lineface.assign(m->lineface, m->lineface + m->nlineelem);
    TBS_TRY(request_sizes(*r, s->prec, s->ctx->nvar, &r->persist, &r->footprint));
    *out = std::move(r);
    return TBS_OK;
}

}  // namespace

extern "C" {

int tbs_server_create(tbs_ctx *ctx, size_t hbm_budget, int precision, double rtol, int max_sweeps,
                      int sweeps_per_step, tbs_server **out) {
    TBS_TRY(tbs::use_ctx(ctx));
    TBS_REQUIRE(out && (precision == TBS_FP32 || precision == TBS_FP64) && rtol >= 0 && max_sweeps >= 1 &&
                    max_sweeps < 4096 && sweeps_per_step >= 1, TBS_ERR_ARG, "bad server arguments");
    *out = nullptr;
    std::unique_ptr<tbs_server, int (*)(tbs_server *)> s(new tbs_server(), tbs_server_destroy);
    s->ctx = ctx;
    TBS_TRY(tbs_ctx_create(ctx->device, ctx->nvar, &s->adm));
    TBS_TRY(tbs_ctx_create(ctx->device, ctx->nvar, &s->ret));
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        TBS_HIP(hipMemGetInfo(&fr, &tot));
        hbm_budget = fr / 10 * 8;
    }
    s->budget = hbm_budget / kAlign * kAlign;
    s->prec = precision;
    s->rtol = rtol;
    s->max_sweeps = max_sweeps;
    s->k = sweeps_per_step;
    // the arena at its full size now, so that no allocation (which waits
    // for the device) falls inside a request's latency; if the device
    // refuses that much at once, smaller first segments, and the rest grows
    // on demand
    for (size_t want = s->budget; want >= std::min(kSegment, s->budget) && want > 0; want /= 2)
        if (grow_arena(s.get(), want) == TBS_OK) break;
    hipGetLastError();
    // the rest of what serving would otherwise allocate on first use, inside
    // some request's latency: pinned result buffers (two of every size class
    // up to 32 MiB, a 460,000-element result at nvar 9), the admission
    // context's table staging ring and device stage, and the batch's tables
    TBS_TRY(s->pinned.prefill((size_t)32 << 20, 2));
    for (int i = 0; i < tbs_ctx::kStages; ++i) {
        char *h = nullptr;
        TBS_TRY(tbs::stage_acquire(s->adm, (size_t)32 << 20, &h));
        TBS_TRY(tbs::stage_release(s->adm));
    }
    TBS_TRY(tbs::dev_stage(s->adm, (size_t)1 << 20));
    TBS_TRY(tbs::batch_create_reserved(ctx, ctx->nvar, precision, 256, (size_t)1 << 16, (size_t)1 << 18, &s->batch));
    hipStream_t aux = nullptr;
    TBS_TRY(tbs::aux_stream(s->adm, &aux));   // the factorisation's packing stream
    TBS_TRY(tbs::aux_stream(ctx, &aux));      // the batch's second half
    // copies of every size a request moves (the runtime picks its copy path
    // by size, and a path's first use is slow): host -> device on the batch
    // and admission streams, device -> host on the retirement stream, into the
    // arena and pinned buffers of the pool
    if (!s->seg_base.empty()) {
        double *hbuf = nullptr;
        size_t cls = 0;
        const size_t big = std::min<size_t>((size_t)32 << 20, s->seg_len[0]);
        TBS_TRY(s->pinned.get(big, &hbuf, &cls));
        for (size_t b = 4096; b <= big; b <<= 2) {
            TBS_HIP(hipMemcpyAsync(s->seg_base[0], hbuf, b, hipMemcpyHostToDevice, ctx->stream));
            TBS_HIP(hipMemcpyAsync(s->seg_base[0], hbuf, b, hipMemcpyHostToDevice, s->adm->stream));
            TBS_HIP(hipMemcpyAsync(hbuf, s->seg_base[0], b, hipMemcpyDeviceToHost, s->ret->stream));
            TBS_HIP(hipDeviceSynchronize());
        }
        s->pinned.put(hbuf, cls);
    }
    TBS_HIP(hipDeviceSynchronize());
    s->t0 = clk::now();
    s->worker = std::thread(admission_loop, s.get());
    // one small request through the whole pipeline (admission thread, batch,
    // retirement), so that no first-use cost of a stream, kernel or buffer
    // lands on a caller's request; then the counters start from zero
    {
        tbs_server_request rq{};
        rq.NL = 2;
        rq.NK = 64;
        rq.seed = 1;
        rq.shift = 4.0 * ctx->nvar;
        long id = 0;
        size_t need = 0;
        TBS_TRY(tbs_server_footprint(s.get(), &rq, &need));
        bool done = need > s->budget;   // (a budget too small for it: no warm-up)
        if (!done) TBS_TRY(tbs_server_submit(s.get(), &rq, &id));
        // (bounded by time, not by steps: in a fresh process the request's
        // set-up includes the first launch of every kernel it uses, and an
        // idle step takes microseconds)
        const auto t_end = clk::now() + std::chrono::seconds(120);
        while (!done && clk::now() < t_end) {
            TBS_TRY(tbs_server_step(s.get(), nullptr, nullptr, nullptr));
            done = s->finished.count(id) != 0;
            if (!done) std::this_thread::yield();
        }
        TBS_REQUIRE(done, TBS_ERR_STATE, "server warm-up request did not finish within 120 s");
        if (id) TBS_TRY(tbs_server_release(s.get(), id));
        s->completed_unpolled.clear();
        s->next_id = 1;
        for (double &t : s->t_phase) t = 0.0;
        for (long &c : s->count) c = 0;
        s->ns_plan = 0;
        s->ns_setup = 0;
        s->t0 = clk::now();
    }
    *out = s.release();
    return TBS_OK;
}

int tbs_server_destroy(tbs_server *s) {
    if (!s) return TBS_OK;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv.notify_all();
        s->worker.join();   // finishes the set-ups already handed over
    }
    hipSetDevice(s->ctx->device);
    hipDeviceSynchronize();
    if (s->batch) tbs_batch_destroy(s->batch);
    for (auto &r : s->active) drop_device(s, r.get());
    for (auto &r : s->admitting) drop_device(s, r.get());
    for (auto &r : s->retiring) drop_device(s, r.get());
    for (char *b : s->seg_base) hipFree(b);
    s->finished.clear();
    s->pinned.clear();
    if (s->ret) tbs_ctx_destroy(s->ret);
    if (s->adm) tbs_ctx_destroy(s->adm);
    delete s;
    return TBS_OK;
}

int tbs_server_footprint(tbs_server *s, const tbs_server_request *req, size_t *bytes) {
    TBS_REQUIRE(s && req && bytes, TBS_ERR_ARG, "null argument");
    Request r;
    r.spec = *req;
    TBS_TRY(build_mesh(&r));
    size_t persist = 0;
    return request_sizes(r, s->prec, s->ctx->nvar, &persist, bytes);
}

int tbs_server_footprint_system(tbs_server *s, const tbs_mesh_host *mesh, size_t *bytes) {
    TBS_REQUIRE(s && bytes, TBS_ERR_ARG, "null argument");
    std::unique_ptr<Request> r;
    TBS_TRY(caller_request(s, mesh, &r));
    *bytes = r->footprint;
    return TBS_OK;
}

int tbs_server_submit_system(tbs_server *s, const tbs_mesh_host *mesh, const double *jacD, const double *jacO1,
                             const double *jacO2, const double *rhs, int where, long *id) {
    TBS_REQUIRE(s && id && jacD && rhs, TBS_ERR_ARG, "null argument");
    TBS_REQUIRE(where == TBS_HOST || where == TBS_DEVICE, TBS_ERR_ARG, "where must be TBS_HOST or TBS_DEVICE");
    std::unique_ptr<Request> r;
    TBS_TRY(caller_request(s, mesh, &r));
    TBS_REQUIRE(r->nintface == 0 || (jacO1 && jacO2), TBS_ERR_ARG, "null face blocks");
    TBS_REQUIRE(r->footprint <= s->budget, TBS_ERR_NOMEM, "request needs %zu bytes, budget is %zu", r->footprint,
                s->budget);
    const int nv = s->ctx->nvar;
    const size_t bs = (size_t)nv * nv, nD = bs * r->nelem, nO = bs * (size_t)r->nintface, nr = (size_t)nv * r->nelem;
    r->where = where;
    if (where == TBS_HOST) {
        r->hD.assign(jacD, jacD + nD);
        if (nO) {
            r->hO1.assign(jacO1, jacO1 + nO);
            r->hO2.assign(jacO2, jacO2 + nO);
        }
        r->hrhs.assign(rhs, rhs + nr);
    } else {
        r->uD = jacD;
        r->uO1 = jacO1;
        r->uO2 = jacO2;
        r->urhs = rhs;
    }
    r->id = s->next_id++;
    r->t_submit = s->now();
    *id = r->id;
    s->queue.push_back(std::move(r));
    return TBS_OK;
}

int tbs_server_timeline(tbs_server *s, long id, double *t, int n) {
    TBS_REQUIRE(s && t, TBS_ERR_ARG, "null argument");
    auto it = s->finished.find(id);
    TBS_REQUIRE(it != s->finished.end(), TBS_ERR_ARG, "request %ld is not finished", id);
    const Request *r = it->second.get();
    const double v[6] = {r->t_submit, r->t_admit, r->t_ready, r->t_conv, r->t_retire, r->t_done};
    for (int i = 0; i < n && i < 6; ++i) t[i] = v[i];
    return TBS_OK;
}

int tbs_server_submit(tbs_server *s, const tbs_server_request *req, long *id) {
    TBS_REQUIRE(s && req && id, TBS_ERR_ARG, "null argument");
    std::unique_ptr<Request> r(new Request());
    r->spec = *req;
    TBS_TRY(build_mesh(r.get()));
    TBS_TRY(request_sizes(*r, s->prec, s->ctx->nvar, &r->persist, &r->footprint));
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
    const int nv = s->ctx->nvar;
    s->count[C_STEPS]++;
    double tp = s->now();
    auto lap = [&](int phase) { const double t = s->now(); s->t_phase[phase] += t - tp; tp = t; };

    // 1. states of the running batch (waits for the sweeps of the last step)
    std::vector<int> st(4 * s->active.size());
    if (!s->active.empty()) TBS_TRY(tbs_batch_states(s->batch, st.data()));
    lap(T_WAIT);

    // 2. membership: finished systems leave, completed admissions join
    std::vector<std::unique_ptr<Request>> keep, done;
    for (size_t i = 0; i < s->active.size(); ++i) {
        if (st[4 * i + 1]) {
            keep.push_back(std::move(s->active[i]));
        } else {
            s->active[i]->status = st[4 * i + 2];
            s->active[i]->t_conv = s->now();
            s->active[i]->nsweeps = std::max(std::min(st[4 * i], 4096) - 1, 0);
            done.push_back(std::move(s->active[i]));
        }
    }
    bool changed = !done.empty();
    for (auto it = s->admitting.begin(); it != s->admitting.end();) {
        Request *r = it->get();
        const int su = r->setup.load(std::memory_order_acquire);
        if (su < 0) {   // set-up failed: the admission thread is done with it
            std::unique_ptr<Request> own = std::move(*it);
            it = s->admitting.erase(it);
            s->used -= own->slab_len;
            drop_device(s, own.get());
            own->status = -1;
            s->count[C_ADMITTED]--;
            finish(s, std::move(own));
            continue;
        }
        if (su == 0 || hipEventQuery(r->ready) != hipSuccess) { ++it; continue; }
        // set-up complete: the temporaries of the slab go back to the arena,
        // and a caller's host copies are no longer needed
        r->t_ready = s->now();
        std::vector<double>().swap(r->hD);
        std::vector<double>().swap(r->hO1);
        std::vector<double>().swap(r->hO2);
        std::vector<double>().swap(r->hrhs);
        s->arena.release(r->slab_off + r->persist, r->slab_len - r->persist);
        s->used -= r->slab_len - r->persist;
        r->slab_len = r->persist;
        keep.push_back(std::move(*it));
        it = s->admitting.erase(it);
        changed = true;
    }
    hipGetLastError();   // hipEventQuery reports "not ready" as an error code
    s->active.swap(keep);
    if (changed) {
        std::vector<tbs_plan *> plans;
        for (auto &r : s->active) plans.push_back(r->plan);
        TBS_TRY(tbs_batch_set_plans(s->batch, plans.data(), (int)plans.size()));
        s->count[C_REBUILDS]++;
    }
    lap(T_MEMBERS);

    // 3. sweeps (asynchronous)
    if (!s->active.empty()) TBS_TRY(tbs_batch_launch_sweeps(s->batch, s->k, s->max_sweeps, s->rtol));
    lap(T_LAUNCH);

    // 4. retire on the retirement stream (no queued kernel touches these
    //    systems), without waiting: a request finishes at the first step that
    //    finds its copies complete, or at once when nothing else is in flight
    if (!done.empty()) {
        hipStream_t rs = s->ret->stream;
        for (auto &r : done) {
            r->nU = (size_t)nv * r->nelem;
            TBS_TRY(s->pinned.get(r->nU * 8, &r->hU, &r->hU_cls));
            r->nH = (size_t)r->nsweeps + 1;
            TBS_TRY(s->pinned.get(r->nH * 8, &r->hH, &r->hH_cls));
            TBS_TRY(tbs::plan_scatter_U(r->plan, rs, r->d_U));
            TBS_HIP(hipMemcpyAsync(r->hU, r->d_U, r->nU * 8, hipMemcpyDeviceToHost, rs));
            TBS_HIP(hipMemcpyAsync(r->hH, tbs::plan_hist(r->plan), r->nH * 8,
                                   hipMemcpyDeviceToHost, rs));
            TBS_HIP(hipEventRecord(r->ready, rs));
            r->t_retire = s->now();
            s->retiring.push_back(std::move(r));
        }
    }
    const bool idle = s->active.empty() && s->admitting.empty() && s->queue.empty();
    TBS_TRY(reap(s, idle));
    lap(T_RETIRE);

    // 5. admission: FIFO; later requests may pass a head that does not fit
    //    until the head has been passed over kMaxSkips times.  An admitted
    //    request's set-up goes to the admission thread.
    std::vector<Request *> fresh;
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Request *r = it->get();
        const bool is_head = it == s->queue.begin();
        bool fits = s->used + r->footprint <= s->budget;
        if (fits && !s->arena.alloc(r->footprint, &r->slab_off)) {
            fits = s->arena_bytes + r->footprint <= s->budget && grow_arena(s, r->footprint) == TBS_OK &&
                   s->arena.alloc(r->footprint, &r->slab_off);
        }
        if (!fits) {
            if (is_head && ++r->skipped > kMaxSkips) break;
            ++it;
            continue;
        }
        r->slab_len = r->footprint;
        r->slab_ptr = s->addr(r->slab_off);
        s->used += r->footprint;
        r->t_admit = s->now();
        s->count[C_ADMITTED]++;
        fresh.push_back(r);
        s->admitting.push_back(std::move(*it));
        it = s->queue.erase(it);
    }
    if (!fresh.empty()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->handoff.insert(s->handoff.end(), fresh.begin(), fresh.end());
        }
        s->cv.notify_one();
    }
    lap(T_ADMIT);

    if (nactive) *nactive = (int)(s->active.size() + s->retiring.size());
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
        TBS_REQUIRE(ucap >= r->nU, TBS_ERR_ARG, "U capacity %zu < %zu", ucap, r->nU);
        if (r->nU) std::memcpy(U, r->hU, r->nU * 8);
    }
    if (hist && hcap > 0 && r->nH) std::memcpy(hist, r->hH, std::min<size_t>(hcap, r->nH) * 8);
    if (nsweeps) *nsweeps = r->nsweeps;
    if (status) *status = r->status;
    if (r->status == -1 && !r->err.empty()) tbs::set_error("admission failed: %s", r->err.c_str());
    if (times) {
        times[0] = r->t_submit;
        times[1] = r->t_admit;
        times[2] = r->t_done;
    }
    return TBS_OK;
}

int tbs_server_release(tbs_server *s, long id) {
    TBS_REQUIRE(s, TBS_ERR_ARG, "null server");
    auto it = s->finished.find(id);
    if (it != s->finished.end()) {
        s->pinned.put(it->second->hU, it->second->hU_cls);
        s->pinned.put(it->second->hH, it->second->hH_cls);
        s->finished.erase(it);
    }
    return TBS_OK;
}

int tbs_server_info(tbs_server *s, size_t *budget, size_t *used, double *now) {
    TBS_REQUIRE(s, TBS_ERR_ARG, "null server");
    if (budget) *budget = s->budget;
    if (used) *used = s->used;
    if (now) *now = s->now();
    return TBS_OK;
}

int tbs_server_stats(tbs_server *s, double *seconds, int nseconds, long *counts, int ncounts) {
    TBS_REQUIRE(s, TBS_ERR_ARG, "null server");
    s->t_phase[T_PLAN] = 1e-9 * (double)s->ns_plan.load();
    s->t_phase[T_SETUP] = 1e-9 * (double)s->ns_setup.load();
    for (int i = 0; seconds && i < nseconds && i < T_NPHASE; ++i) seconds[i] = s->t_phase[i];
    for (int i = 0; counts && i < ncounts && i < C_NCOUNT; ++i) counts[i] = s->count[i];
    return TBS_OK;
}

}  // extern "C"
