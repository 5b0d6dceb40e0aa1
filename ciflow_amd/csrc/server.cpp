// Continuously batched ground-state solve server.
//
// The reference issues one rotated-integral solve after another from its
// orbital optimisers (OrbitalOptimization::update_cim -> CIMethod::set_ham ->
// build_ci -> solve, SimulatedAnnealing.cpp:83-87; LocalMinimizer.cpp:73-235),
// rebuilding the CRS and freeing it after every solve (CIMethod.cpp:117-126,
// 262).  Here requests enter a queue at any time; `slots` worker threads, each
// with its own device context and stream, take the next request whose HBM
// footprint fits the remaining budget (admission), run rotate -> strings ->
// CRS build -> Lanczos on their stream, and publish the result.  Kernels of
// different requests overlap on the device; a request's own kernels are
// deterministic (fixed-order reductions), so its result does not depend on
// what else is in flight.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ciflow_amd.h"
#include "common.h"
#include "davbatch.h"
#include "kernels.h"
#include "spmv_ell.h"

namespace {
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Job {
    cf_request req;
    std::vector<double> h, V, U;
    uint64_t footprint = 0;
    double t_submit = 0.0;
};
}  // namespace

struct cf_server {
    int device = 0;
    uint64_t budget = 0, in_use = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job> pending;
    std::deque<cf_result> done;
    std::vector<std::thread> workers;
    int running = 0;
    uint64_t submitted = 0, finished = 0;
    bool stop = false;
    std::string first_error;
};

namespace {
// Requests the batched one-launch DOCI solver takes (cf_solve_doci_batch):
// the server's own choice (CF_SOLVER_AUTO), DOCI, a space of at most
// kBatchN determinants.
constexpr uint64_t kBatchN = 13312;
constexpr size_t kBatchMax = 256;  // requests per batch (one workgroup each)

bool batchable(const cf_request& r) {
    return r.kind == CF_DOCI && r.method == CF_SOLVER_AUTO && r.nup == r.ndown && r.nup >= 1 && r.nup < r.L &&
           r.L <= 32 && cf::binom_host(r.L, r.nup) <= kBatchN;
}

// Requests the batched Davidson solver takes (cf_solve_batch): the server's
// own choice (CF_SOLVER_AUTO), FCI on the on-the-fly sigma or DOCI above the
// one-launch size, of a shape it has kernels for.
constexpr size_t kDavBatchMax = 32;  // requests per lock-step batch

bool dav_batchable(const cf_request& r) {
    return r.method == CF_SOLVER_AUTO && !batchable(r) && cf::bd_accepts(r, nullptr);
}

uint64_t estimate_footprint(const cf_request& r) {
    const int L = r.L;
    const uint64_t da = cf::binom_host(L, r.nup), db = cf::binom_host(L, r.ndown);
    const int ncv = r.ncv > 0 ? r.ncv : 42;
    const uint64_t ints = 3 * (uint64_t)L * L * L * L * 8;
    if (dav_batchable(r)) {
        // basis and images (2 M vectors), x, H x, t, the FP32 operand copy,
        // integrals and the operator tables
        const uint64_t n = r.kind == CF_DOCI ? da : da * db;
        const uint64_t M = (uint64_t)std::min(std::min(ncv, 24), (int)std::min<uint64_t>(n, 24));
        uint64_t ops = n * 8 + (uint64_t)L * L * 12;  // diagonal, pair tables
        if (r.kind == CF_FCI) {
            const uint64_t dense = (da * da + (r.nup == r.ndown ? 0 : db * db)) * 8;
            ops += dense + (uint64_t)L * L * L * L * 12 + (da + db) * (uint64_t)L * L * 12 + (da + db) * 4096;
        }
        return (2 * M + 3) * n * 8 + n * 4 + ops + ints + (16ull << 20);
    }
    if (r.kind == CF_FCI && r.direct) {
        // Krylov basis + dense FP64 string Hamiltonians + FP32 operand + lists
        const uint64_t n = da * db;
        const uint64_t dense = (da * da + (r.nup == r.ndown ? 0 : db * db)) * 8;
        return (uint64_t)(ncv + 3) * n * 8 + dense + n * 4 + (da + db) * 4096 + ints + (64ull << 20);
    }
    if (batchable(r)) return 3 * da * 8 + 2 * ints + (1ull << 20);  // scratch + integrals (tables shared)
    uint64_t n, row;
    if (r.kind == CF_DOCI) {
        n = da;
        row = 1 + (uint64_t)r.nup * (L - r.nup);
    } else {
        n = da * db;
        const uint64_t sa = (uint64_t)r.nup * (L - r.nup), sb = (uint64_t)r.ndown * (L - r.ndown);
        const uint64_t dda = cf::binom_host(r.nup, 2) * cf::binom_host(L - r.nup, 2);
        const uint64_t ddb = cf::binom_host(r.ndown, 2) * cf::binom_host(L - r.ndown, 2);
        row = 1 + sa + sb + sa * sb + dda + ddb;
    }
    const uint64_t bytes_nnz = r.precision == CF_F32 ? 16 : 12;
    uint64_t crs = n * row * bytes_nnz + (n + 1) * 8 + n * 4;
    if (r.kind == CF_DOCI) crs += n * (row - 1) * 4;  // hop columns of the DOCI SpMV
    // FCI spaces without the position-major copy keep a sliced copy of the
    // rows (cf_crs_build, when it fits): the entries again (padding to 4 per
    // row), chunk offsets and the split launches' range sums
    if (r.kind == CF_FCI && !(da >= 400 && db <= cf::kEllMaxDb) && n >= 8192)
        crs += n * (row + 3) * bytes_nnz + (n / 64 + 2) * 8 + (uint64_t)cf::kSellMaxSplit * n * 8;
    const uint64_t krylov = (uint64_t)(ncv + 3) * n * 8;
    return crs + krylov + ints + (64ull << 20);
}

void publish(cf_server* S, const cf_result& res, uint64_t footprint) {
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->in_use -= footprint;
        S->running--;
        S->finished++;
        S->done.push_back(res);
    }
    S->cv_work.notify_all();
    S->cv_done.notify_all();
}

// Returns HBM held past its requests' results (a batch frees it at its end).
void release(cf_server* S, uint64_t bytes) {
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->in_use -= bytes;
    }
    S->cv_work.notify_all();
}

void run_batch(cf_server* S, cf_ctx* ctx, std::vector<Job>& batch) {
    std::vector<cf_request> reqs(batch.size());
    std::vector<cf_result> res(batch.size());
    for (size_t i = 0; i < batch.size(); ++i) {
        reqs[i] = batch[i].req;
        reqs[i].h = batch[i].h.data();
        reqs[i].V = batch[i].V.data();
        reqs[i].U = batch[i].U.empty() ? nullptr : batch[i].U.data();
    }
    const double t0 = now_s();
    const int st = cf_solve_doci_batch(ctx, (int)batch.size(), reqs.data(), res.data());
    const double t1 = now_s();
    for (size_t i = 0; i < batch.size(); ++i) {
        cf_result r = res[i];
        if (st != CF_OK) {
            std::memset(&r, 0, sizeof(r));
            r.tag = batch[i].req.tag;
            r.status = st;
            std::snprintf(r.error, sizeof(r.error), "%s", cf_last_error());
        }
        r.t_submit = batch[i].t_submit;
        r.t_start = t0;
        r.t_end = t1;
        publish(S, r, batch[i].footprint);
    }
}

// A lock-step batch of same-shape Davidson requests (cf_solve_batch's
// solver): the batch reserves HBM for up to kDavBatchMax slots, takes the
// queued requests of its shape that reach the front of the queue as slots
// free up (continuous batching: a request that arrives while its shape is
// being served joins the running batch instead of waiting for a worker; one
// of another shape at the front stops the refill, so the batch drains and
// frees its worker), publishes each result as soon as its request finishes,
// and releases the reservation when it ends.
void run_dav_batch(cf_server* S, cf_ctx* ctx, std::vector<Job>& first) {
    const uint64_t key = cf::bd_class_key(first[0].req);
    uint64_t fp = 0, held = 0;
    for (const Job& j : first) {
        fp = std::max(fp, j.footprint);
        held += j.footprint;
    }
    int cap = (int)first.size();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        while (cap < (int)kDavBatchMax && S->in_use + fp <= S->budget) {
            S->in_use += fp;
            held += fp;
            ++cap;
        }
    }
    std::deque<Job> jobs;  // by id; stable addresses for the arrays the batch reads
    std::vector<double> t_in;
    std::vector<char> out;
    auto request_of = [](const Job& j) {
        cf_request r = j.req;
        r.h = j.h.data();
        r.V = j.V.data();
        r.U = j.U.empty() ? nullptr : j.U.data();
        return r;
    };
    const double t0 = now_s();
    std::vector<std::pair<int, cf_request>> in;
    for (Job& j : first) {
        jobs.push_back(std::move(j));
        t_in.push_back(t0);
        out.push_back(0);
        in.emplace_back((int)jobs.size() - 1, request_of(jobs.back()));
    }
    cf::BatchFeed feed;
    feed.more = [&](int max, std::vector<std::pair<int, cf_request>>& more) {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return;
        const double t = now_s();
        while ((int)more.size() < max && !S->pending.empty()) {
            Job& f = S->pending.front();
            if (!dav_batchable(f.req) || cf::bd_class_key(f.req) != key || f.footprint > fp) break;
            jobs.push_back(std::move(f));
            S->pending.pop_front();
            S->running++;
            t_in.push_back(t);
            out.push_back(0);
            more.emplace_back((int)jobs.size() - 1, request_of(jobs.back()));
        }
    };
    feed.done = [&](int id, const cf_result& r0) {
        cf_result r = r0;
        r.t_submit = jobs[id].t_submit;
        r.t_start = t_in[id];
        r.t_end = now_s();
        out[id] = 1;
        publish(S, r, 0);
    };
    const int st = cf::solve_batch_feed(ctx, cap, in, feed);
    const double t1 = now_s();
    for (size_t i = 0; i < jobs.size(); ++i) {
        if (out[i]) continue;  // published when it finished
        cf_result r;
        std::memset(&r, 0, sizeof(r));
        r.tag = jobs[i].req.tag;
        r.status = st != CF_OK ? st : CF_ERR_STATE;
        std::snprintf(r.error, sizeof(r.error), "%s", st != CF_OK ? cf_last_error() : "batch ended without a result");
        r.t_submit = jobs[i].t_submit;
        r.t_start = t_in[i];
        r.t_end = t1;
        publish(S, r, 0);
    }
    release(S, held);
}

void worker_main(cf_server* S) {
    cf_ctx* ctx = nullptr;
    if (cf_ctx_create(S->device, &ctx) != CF_OK) {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->first_error.empty()) S->first_error = cf_last_error();
        S->stop = true;
        S->cv_work.notify_all();
        S->cv_done.notify_all();
        return;
    }
    for (;;) {
        Job job;
        std::vector<Job> batch, dav;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop) {
                    lk.unlock();
                    cf_ctx_destroy(ctx);
                    return;
                }
                // first queued job that fits (FIFO among those that fit; a job
                // larger than the whole budget runs alone)
                auto it = S->pending.begin();
                for (; it != S->pending.end(); ++it)
                    if (S->in_use + it->footprint <= S->budget || (S->running == 0 && S->in_use == 0)) break;
                if (it != S->pending.end()) {
                    job = std::move(*it);
                    S->pending.erase(it);
                    S->in_use += job.footprint;
                    S->running++;
                    if (dav_batchable(job.req)) {
                        // with it, queued requests of the same shape that fit
                        const uint64_t key = cf::bd_class_key(job.req);
                        dav.push_back(std::move(job));
                        for (auto jt = S->pending.begin(); jt != S->pending.end() && dav.size() < kDavBatchMax;) {
                            if (dav_batchable(jt->req) && cf::bd_class_key(jt->req) == key &&
                                S->in_use + jt->footprint <= S->budget) {
                                S->in_use += jt->footprint;
                                S->running++;
                                dav.push_back(std::move(*jt));
                                jt = S->pending.erase(jt);
                            } else {
                                ++jt;
                            }
                        }
                    } else if (batchable(job.req)) {
                        // with it, every other queued small DOCI request of the same
                        // shape that fits (other shapes go to other workers, so a
                        // batch of small spaces does not wait for larger ones)
                        const int L0 = job.req.L, np0 = job.req.nup;
                        batch.push_back(std::move(job));
                        for (auto jt = S->pending.begin(); jt != S->pending.end() && batch.size() < kBatchMax;) {
                            if (batchable(jt->req) && jt->req.L == L0 && jt->req.nup == np0 &&
                                S->in_use + jt->footprint <= S->budget) {
                                S->in_use += jt->footprint;
                                S->running++;
                                batch.push_back(std::move(*jt));
                                jt = S->pending.erase(jt);
                            } else {
                                ++jt;
                            }
                        }
                    }
                    break;
                }
                S->cv_work.wait(lk);
            }
        }
        if (!batch.empty()) {
            run_batch(S, ctx, batch);
            continue;
        }
        if (!dav.empty()) {
            run_dav_batch(S, ctx, dav);
            continue;
        }
        cf_result res;
        std::memset(&res, 0, sizeof(res));
        res.tag = job.req.tag;
        res.t_submit = job.t_submit;
        res.t_start = now_s();
        res.n = 0;
        const cf_request& r = job.req;
        cf_ham *ham = nullptr, *hr = nullptr;
        cf_space* sp = nullptr;
        cf_crs* m = nullptr;
        cf_sigma* sg = nullptr;
        const bool direct = r.kind == CF_FCI && r.direct;
        int st = cf_ham_create(ctx, r.L, r.nup, r.ndown, job.h.data(), job.V.data(), r.econst, &ham);
        if (st == CF_OK && !job.U.empty()) st = cf_ham_rotate(ham, job.U.data(), &hr);
        if (st == CF_OK) st = cf_space_create(hr ? hr : ham, r.kind, &sp);
        if (st == CF_OK) st = direct ? cf_sigma_create(sp, r.precision, &sg) : cf_crs_build(sp, r.precision, 1e-14, &m);
        if (st == CF_OK) {
            cf_solve_opts o;
            cf_solve_opts_default(&o);
            if (r.ncv > 0) o.ncv = r.ncv;
            if (r.tol > 0) o.tol = r.tol;
            o.precision = r.precision;
            o.check_every = r.check_every > 0 ? r.check_every : 0;
            o.tol_energy = r.tol_energy > 0.0 ? r.tol_energy : 0.0;
            o.method = r.method;
            if (r.method == CF_SOLVER_AUTO) {
                // measured per class on MI355X (8 workers, solves/s, Davidson
                // against Lanczos): FCI L=8-11 609 / 568, 433 / 384, 255 / 240,
                // 81-90 / 68; DOCI L=14-20 1004 / 966 ... 246 / 218; DOCI L=12
                // (924 determinants, launch-bound) 1066 / 1138
                uint64_t n = 0;
                cf_space_dims(sp, &n, nullptr, nullptr);
                o.method = (r.kind == CF_FCI || n >= 2000) ? CF_DAVIDSON : CF_LANCZOS;
            }
            cf_solve_info info;
            st = direct ? cf_solve_direct(sg, &o, nullptr, &info) : cf_solve(m, &o, nullptr, &info);
            if (st == CF_OK) {
                res.energy = info.energy;
                res.eigval = info.eigval;
                res.resid = info.resid;
                res.n_matvec = info.n_matvec;
                res.n_matvec_f32 = info.n_matvec_f32;
                res.solve_status = info.status;
            }
            uint64_t n = 0;
            cf_space_dims(sp, &n, nullptr, nullptr);
            res.n = n;
        }
        res.status = st;
        if (st != CF_OK) {
            std::snprintf(res.error, sizeof(res.error), "%s", cf_last_error());
        }
        cf_crs_destroy(m);
        cf_sigma_destroy(sg);
        cf_space_destroy(sp);
        cf_ham_destroy(hr);
        cf_ham_destroy(ham);
        cf_ctx_synchronize(ctx);
        res.t_end = now_s();
        publish(S, res, job.footprint);
    }
}
}  // namespace

#define SRV_BEGIN try {
#define SRV_END                                       \
    }                                                 \
    catch (const cf::Error& e) {                      \
        return cf::set_last_error(e.code, e.what());  \
    }                                                 \
    catch (const std::exception& e) {                 \
        return cf::set_last_error(CF_ERR_HIP, e.what()); \
    }                                                 \
    return CF_OK;

extern "C" {

// CF_PREC_AUTO: FP32 values only where the request's route stores matrix
// values (a stored CRS); the one-launch DOCI batch, the lock-step Davidson
// batch and the on-the-fly sigma read none, and FP64 is as fast or faster
// there (measured per class on MI355X: FP64 lock-step batches 0-21% more
// solves/s, the FP32 phase ending in 3-12 extra FP64 polish products).
static cf_request resolve_precision(const cf_request& r) {
    cf_request q = r;
    if (q.precision != CF_PREC_AUTO) return q;
    q.precision = CF_F64;
    const bool stored = !(batchable(q) || dav_batchable(q) || (q.kind == CF_FCI && q.direct));
    if (stored) q.precision = CF_F32;
    return q;
}

int cf_request_footprint(const cf_request* r, uint64_t* bytes) {
    SRV_BEGIN
    CF_REQUIRE(r && bytes, "NULL argument");
    CF_REQUIRE(r->L >= 1 && r->L <= cf::kMaxL && r->nup >= 0 && r->nup <= r->L && r->ndown >= 0 &&
                   r->ndown <= r->L,
               "bad request");
    *bytes = estimate_footprint(resolve_precision(*r));
    SRV_END
}

int cf_server_create(int device, int slots, double hbm_budget_bytes, cf_server** out) {
    SRV_BEGIN
    CF_REQUIRE(out && slots >= 1 && slots <= 64, "bad arguments");
    std::unique_ptr<cf_server> S(new cf_server());
    S->device = device;
    if (hbm_budget_bytes <= 0) {
        size_t fr = 0, tot = 0;
        CF_HIP(hipSetDevice(device));
        CF_HIP(hipMemGetInfo(&fr, &tot));
        S->budget = (uint64_t)(0.8 * (double)fr);
    } else {
        S->budget = (uint64_t)hbm_budget_bytes;
    }
    for (int i = 0; i < slots; ++i) S->workers.emplace_back(worker_main, S.get());
    *out = S.release();
    SRV_END
}

int cf_server_submit(cf_server* S, const cf_request* r) {
    SRV_BEGIN
    CF_REQUIRE(S && r && r->h && r->V, "NULL argument");
    CF_REQUIRE(r->L >= 1 && r->L <= cf::kMaxL && (r->kind == CF_FCI || r->kind == CF_DOCI), "bad request");
    CF_REQUIRE(r->precision == CF_F64 || r->precision == CF_F32 || r->precision == CF_PREC_AUTO, "bad precision");
    CF_REQUIRE(r->nup >= 0 && r->nup <= r->L && r->ndown >= 0 && r->ndown <= r->L, "electron counts out of range");
    Job j;
    j.req = resolve_precision(*r);
    const size_t L2 = (size_t)r->L * r->L;
    j.h.assign(r->h, r->h + L2);
    j.V.assign(r->V, r->V + L2 * L2);
    if (r->U) j.U.assign(r->U, r->U + L2);
    j.footprint = estimate_footprint(j.req);
    j.t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        CF_REQUIRE(!S->stop, S->first_error.empty() ? "server stopped" : S->first_error);
        S->pending.push_back(std::move(j));
        S->submitted++;
    }
    S->cv_work.notify_one();
    SRV_END
}

int cf_server_poll(cf_server* S, cf_result* out, int max, int* count) {
    SRV_BEGIN
    CF_REQUIRE(S && count && (max == 0 || out), "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    int c = 0;
    while (c < max && !S->done.empty()) {
        out[c++] = S->done.front();
        S->done.pop_front();
    }
    *count = c;
    SRV_END
}

int cf_server_drain(cf_server* S, double timeout_s) {
    SRV_BEGIN
    CF_REQUIRE(S, "NULL argument");
    std::unique_lock<std::mutex> lk(S->mu);
    const auto pred = [S] { return S->stop || S->finished == S->submitted; };
    if (timeout_s > 0) {
        if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
            throw cf::Error(CF_ERR_STATE, "drain timed out");
    } else {
        S->cv_done.wait(lk, pred);
    }
    CF_REQUIRE(S->first_error.empty(), S->first_error);
    SRV_END
}

int cf_server_stats(cf_server* S, uint64_t* submitted, uint64_t* finished, uint64_t* in_use, uint64_t* budget) {
    SRV_BEGIN
    CF_REQUIRE(S, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    if (submitted) *submitted = S->submitted;
    if (finished) *finished = S->finished;
    if (in_use) *in_use = S->in_use;
    if (budget) *budget = S->budget;
    SRV_END
}

int cf_server_destroy(cf_server* S) {
    SRV_BEGIN
    if (!S) return CF_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    delete S;
    SRV_END
}

}  // extern "C"
