// Batched Davidson solves: the served stream's FCI (on-the-fly sigma) and
// larger DOCI requests of one shape advanced in lock step, one launch per
// step for up to 32 requests.  The reference solves each request alone
// (OrbitalOptimization::update_cim -> set_ham -> build_ci -> solve,
// SimulatedAnnealing.cpp:83-87; CIMethod::solve, CIMethod.cpp:241-268); here
// the per-request chain of short launches -- the served stream's bound --
// becomes one chain per batch.
//
// Each lane runs the algorithm of the single-request Davidson solver
// (capi.cpp, run_davidson): an FP32-operand phase from the seeded start
// vector (8 Krylov steps, then diagonal-preconditioned corrections) that
// stops on tol_f32, the energy estimate or a stall of 24 steps, then an FP64
// polish from its Ritz vector that stops on tol or the energy estimate; a
// basis of M = min(ncv, 24) vectors restarted to its 4 lowest Ritz vectors.
// Lanes advance independently inside the batch (each has its own basis size,
// phase and stops; a finished lane idles), the host keeps each lane's
// projected matrix, and a lane's device arithmetic depends on n alone, so a
// request's result is the same in any batch and alone.
#include "davbatch.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <vector>

#include "capi_types.h"
#include "kernels.h"
#include "lanczos.h"
#include "sigma.h"
#include "spmv_doci.h"
#include "transform.h"

using namespace cf;

namespace {

constexpr double kDociZero = 1e-14;  // elements the served DOCI build drops (|v| <= 1e-14)

int basis_size(const cf_request& r, uint64_t n) {
    const int ncv = r.ncv > 0 ? r.ncv : 42;
    return (int)std::min<uint64_t>((uint64_t)std::min(ncv, kDavMaxBasis), n);
}

uint64_t space_dim(const cf_request& r) {
    return r.kind == CF_DOCI ? binom_host(r.L, r.nup) : binom_host(r.L, r.nup) * binom_host(r.L, r.ndown);
}

// The alpha-beta sigma shape of an FCI request (lists only, no device data).
SigmaArgs fci_shape(const cf_request& r) {
    SigmaArgs a;
    a.da = binom_host(r.L, r.nup);
    a.db = binom_host(r.L, r.ndown);
    a.L = r.L;
    a.Sa = 1 + r.nup * (r.L - r.nup);
    a.Sb = 1 + r.ndown * (r.L - r.ndown);
    return a;
}

struct Lane {
    int id = -1;  // the caller's id of the request in this slot; -1: free
    cf_request req{};
    std::unique_ptr<cf_ham> ham;
    std::unique_ptr<cf_space> sp;
    DBuf<double> Ha, Hb, vpair64, vself64, vk64;
    DBuf<float> vpair32, vself32, vk32;
    SigmaArgs args;
    // Davidson state (the fields of run_davidson / Davidson::run)
    int phase = 2;  // 0 FP32 operator, 1 FP64 operator, 2 finished
    bool start = true;
    int start_mode = BD_START_SEED;
    int m = 0, it = 0, krylov = 8, stall = 200;
    double tol = 1e-12, tol_energy = 0.0;
    double best = 0.0, theta_best = 0.0;
    int since_best = 0;
    std::vector<double> G;
    int napply = 0, nmv32 = 0, nrestart = 0;
    int mk = 0;
    double th0 = 0.0, th1 = 0.0, res = 0.0;
    bool have_th1 = false;
    double theta_out = 0.0, res_out = 0.0;
    int code = 0;
    bool failed = false;
    std::string error;
};

struct Batch {
    cf_ctx* c;
    const BatchFeed* feed;
    int kind, L, nup, ndown, M, keep, cap;
    uint64_t n, da, db;
    double tol_f32;
    uint64_t seed;
    int maxit;
    std::vector<Lane> lanes;  // cap slots
    std::vector<BdOpLane> hop;
    std::vector<SigmaArgs> hargs;
    DBuf<double> V, S, x, t, partial, parts, dg, Y;
    DBuf<float> x32;
    DBuf<BdStep> dstep;
    DBuf<BdOpLane> dop;
    DBuf<SigmaArgs> dargs;
    cf_ctx::DociBlockTables* blk = nullptr;
    BdView view;
    // page-locked staging
    BdStep* hstep = nullptr;
    double* hdg = nullptr;
    double* hr1 = nullptr;
    double* hY = nullptr;

    void alloc();
    void add(const std::vector<std::pair<int, cf_request>>& in);
    void run();
    void finish();
    void end_phase(Lane& q, int code);
    void column(Lane& q, const double* dg);
    void product(bool any32, bool any64);
};

// The batch's vectors for `cap` slots (a slot's vectors at a fixed stride).
void Batch::alloc() {
    hipStream_t st = c->stream;
    lanes.resize(cap);
    hop.assign(cap, BdOpLane{});
    hargs.assign(cap, SigmaArgs{});
    if (kind == CF_DOCI) {
        auto key = std::make_pair(L, nup);
        if (!c->doci_blk.count(key)) make_doci_block_tables(c, L, nup, doci_block_w(L), c->doci_blk[key]);
        blk = &c->doci_blk[key];
        blk->args.zero = kDociZero;
    }
    const size_t Bn = (size_t)cap * n;
    V.alloc(Bn * M);
    S.alloc(Bn * M);
    x.alloc(Bn);
    t.alloc(Bn);
    if (kind == CF_FCI) x32.alloc(Bn);
    partial.alloc((size_t)cap * kBdPartial);
    parts.alloc((size_t)cap * kBdParts);
    dg.alloc((size_t)cap * 64);
    Y.alloc((size_t)cap * kDavMaxBasis * kDavMaxBasis);
    dstep.alloc(cap);
    dop.alloc(cap);
    if (kind == CF_FCI) dargs.alloc(cap);
    view.n = n;
    view.M = M;
    view.Ga = fused_axpy_blocks(n);
    view.Gd = fused_dot_blocks(n);
    view.V = V.p;
    view.S = S.p;
    view.x = x.p;
    view.t = t.p;
    view.x32 = x32.p;
    view.partial = partial.p;
    view.parts = parts.p;
    view.dg = dg.p;
    view.Y = Y.p;
    view.step = dstep.p;
    view.op = dop.p;
    // staging: steps, dg rows, second-pair residual parts, restart rotations
    const size_t bytes = cap * (sizeof(BdStep) + 64 * sizeof(double) + 256 * sizeof(double) +
                                kDavMaxBasis * kDavMaxBasis * sizeof(double));
    char* p = (char*)c->batch_pinned(bytes);
    hstep = (BdStep*)p;
    p += cap * sizeof(BdStep);
    hdg = (double*)p;
    p += cap * 64 * sizeof(double);
    hr1 = (double*)p;
    p += cap * 256 * sizeof(double);
    hY = (double*)p;
    (void)st;
}

// Requests into free slots: rotated integrals (one batched FP64-MFMA
// transform), strings with their H-dependent tables, diagonal, operator
// tables, and a fresh solver state.
void Batch::add(const std::vector<std::pair<int, cf_request>>& in) {
    if (in.empty()) return;
    hipStream_t st = c->stream;
    const size_t L2 = (size_t)L * L, L3 = L2 * L, L4 = L3 * L;
    std::vector<int> slot;
    for (int b = 0; b < cap && slot.size() < in.size(); ++b)
        if (lanes[b].id < 0) slot.push_back(b);
    CF_REQUIRE(slot.size() == in.size(), "batched Davidson: more requests than free slots");
    const int k = (int)in.size();
    std::vector<double> hh(k * L2), vv(k * L4), uu(k * L2, 0.0);
    for (int i = 0; i < k; ++i) {
        const cf_request& r = in[i].second;
        std::memcpy(&hh[i * L2], r.h, L2 * sizeof(double));
        std::memcpy(&vv[i * L4], r.V, L4 * sizeof(double));
        if (r.U) std::memcpy(&uu[i * L2], r.U, L2 * sizeof(double));
        else
            for (int a = 0; a < L; ++a) uu[i * L2 + a * L + a] = 1.0;
    }
    DBuf<double> dh(k * L2), dV(k * L4), dU(k * L2), oh(k * L2), oV(k * L4), tmp(k * L4);
    CF_HIP(hipMemcpyAsync(dh.p, hh.data(), dh.bytes(), hipMemcpyHostToDevice, st));
    CF_HIP(hipMemcpyAsync(dV.p, vv.data(), dV.bytes(), hipMemcpyHostToDevice, st));
    CF_HIP(hipMemcpyAsync(dU.p, uu.data(), dU.bytes(), hipMemcpyHostToDevice, st));
    rotate_integrals_dev(L, k, dh.p, dV.p, dU.p, oh.p, oV.p, tmp.p, st);
    for (int i = 0; i < k; ++i) {
        const int b = slot[i];
        Lane& q = lanes[b];
        q = Lane();
        q.id = in[i].first;
        q.req = in[i].second;
        const cf_request& r = q.req;
        const bool f32 = r.precision == CF_F32;
        q.ham.reset(new cf_ham());
        cf_ham* H = q.ham.get();
        H->ctx = c;
        H->L = L;
        H->nup = nup;
        H->ndown = ndown;
        H->econst = r.econst;
        H->dh.alloc(L2);
        H->dV.alloc(L4);
        H->dW.alloc(L3);
        H->dJm.alloc(L2);
        CF_HIP(hipMemcpyAsync(H->dh.p, oh.p + i * L2, L2 * sizeof(double), hipMemcpyDeviceToDevice, st));
        CF_HIP(hipMemcpyAsync(H->dV.p, oV.p + i * L4, L4 * sizeof(double), hipMemcpyDeviceToDevice, st));
        launch_derived(L, H->dV.p, H->dW.p, H->dJm.p, st);
        q.sp.reset(new cf_space());
        cf_space* sp = q.sp.get();
        sp->ham = H;
        sp->kind = kind;
        if (kind == CF_DOCI) {
            sp->ta.reset(make_strtab(c, H, nup, false));
            sp->tb = sp->ta.get();
            sp->da = sp->db = sp->ta->d;
        } else {
            sp->ta.reset(make_strtab(c, H, nup, true));
            if (ndown == nup) {
                sp->tb = sp->ta.get();
            } else {
                sp->tb_own.reset(make_strtab(c, H, ndown, true));
                sp->tb = sp->tb_own.get();
            }
            sp->da = sp->ta->d;
            sp->db = sp->tb->d;
        }
        sp->n = n;
        sp->diag.alloc(n);
        fill_space_diag(sp);
        BdOpLane& o = hop[b];
        o = BdOpLane{};
        o.diag = sp->diag.p;
        if (kind == CF_DOCI) {
            DociEllArgs ea;
            ea.n = n;
            ea.L = L;
            ea.np = nup;
            ea.nv = L - nup;
            ea.s = sp->ta->s.p;
            ea.diag = sp->diag.p;
            ea.zero = kDociZero;
            q.vk64.alloc(L2);
            q.vk32.alloc(L2);
            launch_doci_pair_table(ea, H->dev(), q.vk64.p, q.vk32.p, st);
            o.vk64 = q.vk64.p;
            o.vk32 = q.vk32.p;
        } else {
            q.Ha.alloc(da * da);
            launch_dense_string_ham<double>(da, sp->ta->E, sp->ta->full.p, q.Ha.p, st);
            o.Ha = o.Hb = q.Ha.p;
            if (sp->tb != sp->ta.get()) {
                q.Hb.alloc(db * db);
                launch_dense_string_ham<double>(db, sp->tb->E, sp->tb->full.p, q.Hb.p, st);
                o.Hb = q.Hb.p;
            }
            SigmaArgs& a = q.args;
            a.da = da;
            a.db = db;
            a.L = L;
            a.Sa = sp->ta->S;
            a.Sb = sp->tb->S;
            a.sa = sp->ta->s.p;
            a.sb = sp->tb->s.p;
            a.suba = sp->ta->sub.p;
            a.subb = sp->tb->sub.p;
            a.V = H->dV.p;
            q.vpair64.alloc(L4);
            q.vself64.alloc(da * L2);
            launch_sigma_tables<double>(a, q.vpair64.p, q.vself64.p, st);
            a.vpair64 = q.vpair64.p;
            a.vself64 = q.vself64.p;
            if (f32) {
                q.vpair32.alloc(L4);
                q.vself32.alloc(da * L2);
                launch_sigma_tables<float>(a, q.vpair32.p, q.vself32.p, st);
                a.vpair32 = q.vpair32.p;
                a.vself32 = q.vself32.p;
            }
            CF_REQUIRE(sigma_ab3_batch_args(a, ndown, false), "batched sigma: no kernel for this shape");
            hargs[b] = a;
        }
        // solver state: the FP32 phase first for FP32 requests
        q.G.assign((size_t)M * M, 0.0);
        q.tol_energy = r.tol_energy > 0.0 ? r.tol_energy : 0.0;
        q.best = std::numeric_limits<double>::infinity();
        if (f32) {
            q.phase = 0;
            q.tol = tol_f32;
            q.stall = 24;
        } else {
            q.phase = 1;
            q.tol = r.tol > 0 ? r.tol : 1e-12;
            q.stall = 200;
        }
        q.krylov = 8;
        q.start = true;
        q.start_mode = BD_START_SEED;
    }
    CF_HIP(hipMemcpyAsync(dop.p, hop.data(), cap * sizeof(BdOpLane), hipMemcpyHostToDevice, st));
    if (kind == CF_FCI)
        CF_HIP(hipMemcpyAsync(dargs.p, hargs.data(), cap * sizeof(SigmaArgs), hipMemcpyHostToDevice, st));
    CF_HIP(hipStreamSynchronize(st));  // the host staging arrays go out of scope
}

// S_m = H V_m for every active lane: FCI same-spin GEMM then the alpha-beta
// kernel of the lane's precision; DOCI the block-staged product.
void Batch::product(bool any32, bool any64) {
    hipStream_t st = c->stream;
    if (kind == CF_FCI) {
        bd_same_spin(view, cap, da, db, st);
        SigmaBatch sb;
        sb.args = dargs.p;
        sb.step = dstep.p;
        sb.V = V.p;
        sb.S = S.p;
        sb.x32 = x32.p;
        sb.n = n;
        sb.M = M;
        const SigmaArgs* shape = nullptr;
        for (const Lane& q : lanes)
            if (q.id >= 0) shape = &q.args;
        if (any32) launch_sigma_ab3_batch<float>(*shape, ndown, sb, cap, st);
        if (any64) launch_sigma_ab3_batch<double>(*shape, ndown, sb, cap, st);
    } else {
        DociBatch bt;
        bt.step = dstep.p;
        bt.op = dop.p;
        bt.V = V.p;
        bt.S = S.p;
        bt.n = n;
        bt.M = M;
        if (any32) launch_spmv_doci_blk_batch(blk->args, true, bt, cap, st);
        if (any64) launch_spmv_doci_blk_batch(blk->args, false, bt, cap, st);
    }
}

void Batch::end_phase(Lane& q, int code) {
    if (q.phase == 0) {  // FP32 phase over: FP64 polish from its Ritz vector
        const cf_request& r = q.req;
        q.nmv32 = q.napply;
        q.napply = 0;
        q.phase = 1;
        q.start = true;
        q.start_mode = BD_START_X;
        q.it = 0;
        q.krylov = 0;
        q.tol = r.tol > 0 ? r.tol : 1e-12;
        q.stall = 200;
        q.best = std::numeric_limits<double>::infinity();
        q.theta_best = 0.0;
        q.since_best = 0;
        q.m = 0;
        std::fill(q.G.begin(), q.G.end(), 0.0);
    } else {
        q.phase = 2;
        q.code = code;
    }
}

// The checks after the stops (breakdown, iteration limit, stall), then the
// new column of the projected matrix.
void Batch::column(Lane& q, const double* d) {
    if (!(d[63] > 1e-28)) return end_phase(q, 2);
    if (q.it >= maxit) return end_phase(q, 1);
    if (q.res < 0.9 * q.best || q.th0 < q.theta_best - q.best) {
        q.best = q.res;
        q.theta_best = q.th0;
        q.since_best = 0;
    } else if (q.stall > 0 && ++q.since_best >= q.stall) {
        return end_phase(q, 3);
    }
    for (int i = 0; i <= q.m; ++i) q.G[(size_t)i * M + q.m] = q.G[(size_t)q.m * M + i] = d[i];
    ++q.m;
    ++q.it;
}

// Results of the lanes that finished: handed to the caller, slots freed.
void Batch::finish() {
    for (Lane& q : lanes) {
        if (q.id < 0 || q.phase != 2) continue;
        cf_result o;
        std::memset(&o, 0, sizeof(o));
        o.tag = q.req.tag;
        o.n = n;
        if (q.failed) {
            o.status = CF_ERR_NOCONV;
            std::snprintf(o.error, sizeof(o.error), "%s", q.error.c_str());
        } else {
            o.status = CF_OK;
            o.eigval = q.theta_out;
            o.energy = q.theta_out + q.req.econst;
            o.resid = q.res_out;
            o.n_matvec = q.napply;
            o.n_matvec_f32 = q.nmv32;
            o.solve_status = q.code == 0 ? 0 : 1;
        }
        const int id = q.id;
        q = Lane();  // releases the slot's tables (ordered on the stream)
        if (feed->done) feed->done(id, o);
    }
}

void Batch::run() {
    hipStream_t st = c->stream;
    std::vector<double> A, th, Yv;
    std::vector<std::pair<int, cf_request>> more;
    for (;;) {
        if (feed->more) {  // free slots take queued requests of the same shape
            int nfree = 0;
            for (const Lane& q : lanes) nfree += q.id < 0;
            if (nfree) {
                more.clear();
                feed->more(nfree, more);
                add(more);
            }
        }
        bool any = false, any_restart = false, any32 = false, any64 = false;
        for (int b = 0; b < cap; ++b) {
            Lane& q = lanes[b];
            BdStep& s = hstep[b];
            std::memset(&s, 0, sizeof(s));
            if (q.id < 0 || q.phase == 2) continue;
            any = true;
            s.f32 = q.phase == 0 ? 1 : 0;
            any32 |= s.f32 != 0;
            any64 |= s.f32 == 0;
            s.seed = seed;
            if (q.start) {
                s.mode = q.start_mode;
                s.m = 0;
                continue;
            }
            const int m = q.m;
            A.assign((size_t)m * m, 0.0);
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j) A[(size_t)i * m + j] = q.G[(size_t)i * M + j];
            sym_eig(m, A, th, Yv);
            if (m == M) {  // restart on the keep lowest Ritz vectors
                std::memcpy(hY + (size_t)b * kDavMaxBasis * kDavMaxBasis, Yv.data(), sizeof(double) * m * m);
                s.restart = 1;
                s.mr = m;
                s.keep = keep;
                any_restart = true;
                std::fill(q.G.begin(), q.G.end(), 0.0);
                for (int i = 0; i < keep; ++i) q.G[(size_t)i * M + i] = th[i];
                q.m = keep;
                ++q.nrestart;
                for (int i = 0; i < q.m; ++i) {
                    s.ys[i] = i == 0 ? 1.0 : 0.0;
                    s.y1[i] = i == 1 ? 1.0 : 0.0;
                }
            } else {
                for (int i = 0; i < m; ++i) {
                    s.ys[i] = Yv[(size_t)i * m + 0];
                    s.y1[i] = m >= 2 ? Yv[(size_t)i * m + 1] : 0.0;
                }
            }
            s.mode = BD_RITZ;
            s.m = q.m;
            s.theta = th[0];
            s.theta1 = th.size() >= 2 ? th[1] : 0.0;
            s.shift = q.it < q.krylov ? 0.0 : 1e-4;
            q.mk = q.m;
            q.th0 = th[0];
            q.have_th1 = th.size() >= 2;
            q.th1 = q.have_th1 ? th[1] : 0.0;
        }
        if (!any) return;
        CF_HIP(hipMemcpyAsync(dstep.p, hstep, cap * sizeof(BdStep), hipMemcpyHostToDevice, st));
        if (any_restart)
            CF_HIP(hipMemcpyAsync(Y.p, hY, (size_t)cap * kDavMaxBasis * kDavMaxBasis * sizeof(double),
                                  hipMemcpyHostToDevice, st));
        bd_form_vectors(view, cap, any_restart, st);
        product(any32, any64);
        bd_new_column(view, cap, st);
        CF_HIP(hipMemcpyAsync(hdg, dg.p, (size_t)cap * 64 * sizeof(double), hipMemcpyDeviceToHost, st));
        CF_HIP(hipStreamSynchronize(st));
        std::vector<int> temple;
        for (int b = 0; b < cap; ++b) {
            Lane& q = lanes[b];
            if (hstep[b].mode == BD_IDLE) continue;
            const double* d = hdg + (size_t)b * 64;
            ++q.napply;
            if (q.start) {
                if (!(d[63] > 1e-28)) {
                    q.failed = true;
                    q.error = "Davidson start vector is zero";
                    q.phase = 2;
                    continue;
                }
                q.G[0] = d[0];
                q.m = 1;
                q.start = false;
                continue;
            }
            q.res = std::sqrt(std::max(0.0, d[62]));
            q.theta_out = q.th0;
            q.res_out = q.res;
            if (q.res <= q.tol * std::fabs(q.th0)) {
                end_phase(q, 0);
                continue;
            }
            // the optimistic gap theta1 - theta0 first, then the safeguarded
            // one with the second pair's residual (one more pass, below)
            if (q.tol_energy > 0.0 && q.mk >= 2 && q.have_th1 && q.res * q.res <= q.tol_energy * (q.th1 - q.th0)) {
                temple.push_back(b);
                continue;
            }
            column(q, d);
        }
        if (!temple.empty()) {
            for (int b : temple) hstep[b].r2 = 1;
            CF_HIP(hipMemcpyAsync(dstep.p, hstep, cap * sizeof(BdStep), hipMemcpyHostToDevice, st));
            bd_resid2(view, cap, st);
            CF_HIP(hipMemcpy2DAsync(hr1, 256 * sizeof(double), parts.p + 1024, kBdParts * sizeof(double),
                                    view.Ga * sizeof(double), cap, hipMemcpyDeviceToHost, st));
            CF_HIP(hipStreamSynchronize(st));
            for (int b : temple) {
                Lane& q = lanes[b];
                double s2 = 0.0;
                for (int i = 0; i < view.Ga; ++i) s2 += hr1[(size_t)b * 256 + i];
                const double r1 = std::sqrt(std::max(0.0, s2));
                if (temple_accept(q.res, q.th0, q.th1, r1, q.tol_energy)) end_phase(q, 0);
                else column(q, hdg + (size_t)b * 64);
            }
        }
        finish();
    }
}

}  // namespace

namespace cf {

bool bd_accepts(const cf_request& r, std::string* why) {
    auto no = [why](const char* w) {
        if (why) *why = w;
        return false;
    };
    if (!r.h || !r.V) return no("NULL integrals");
    if (r.method != CF_SOLVER_AUTO && r.method != CF_DAVIDSON) return no("batched solves run Davidson (method AUTO or DAVIDSON)");
    if (r.precision != CF_F32 && r.precision != CF_F64) return no("precision must be CF_F32 or CF_F64");
    if (r.L < 2 || r.L > 31) return no("batched solves need 2 <= L <= 31");
    if (r.nup < 1 || r.nup >= r.L || r.ndown < 1 || r.ndown >= r.L) return no("batched solves need 1 <= nalpha, nbeta < L");
    if (r.kind == CF_DOCI) {
        if (r.nup != r.ndown) return no("DOCI needs nalpha == nbeta");
        const int w = doci_block_w(r.L);
        if (w < 1 || r.L - w > 20) return no("DOCI shape outside the block-staged product");
        if (binom_host(r.L, r.nup) < 2) return no("space of one determinant");
        return true;
    }
    if (r.kind != CF_FCI) return no("kind must be CF_FCI or CF_DOCI");
    if (!r.direct) return no("batched FCI solves use the on-the-fly sigma (direct = 1)");
    const uint64_t da = binom_host(r.L, r.nup), db = binom_host(r.L, r.ndown);
    if (da * db < 2 || da * db >= (1ull << 32)) return no("FCI dimension out of range");
    SigmaArgs a = fci_shape(r);
    if (!sigma_ab3_batch_args(a, r.ndown, false)) return no("no batched alpha-beta kernel for this shape");
    SigmaArgs a32 = fci_shape(r);
    if (r.precision == CF_F32 && !sigma_ab3_batch_args(a32, r.ndown, true))
        return no("no batched FP32 alpha-beta kernel for this shape");
    return true;
}

uint64_t bd_class_key(const cf_request& r) {
    const int M = basis_size(r, space_dim(r));
    return ((uint64_t)(r.kind & 0xff) << 32) | ((uint64_t)(r.L & 0xff) << 24) | ((uint64_t)(r.nup & 0xff) << 16) |
           ((uint64_t)(r.ndown & 0xff) << 8) | (uint64_t)(M & 0xff);
}

}  // namespace cf

namespace cf {

int solve_batch_feed(cf_ctx* ctx, int capacity, const std::vector<std::pair<int, cf_request>>& first,
                     const BatchFeed& feed) {
    try {
        CF_REQUIRE(ctx && !first.empty(), "batched Davidson: no requests");
        CF_REQUIRE(capacity >= (int)first.size() && capacity <= kBdMaxLanes, "batched Davidson: 1..32 lanes");
        ctx->use();
        cf_solve_opts dflt;
        cf_solve_opts_default(&dflt);
        const cf_request& r0 = first[0].second;
        const uint64_t key = bd_class_key(r0);
        for (const auto& pr : first) {
            std::string why;
            CF_REQUIRE(bd_accepts(pr.second, &why), why);
            CF_REQUIRE(bd_class_key(pr.second) == key, "batched Davidson: requests of different shapes");
        }
        Batch bt;
        bt.c = ctx;
        bt.feed = &feed;
        bt.kind = r0.kind;
        bt.L = r0.L;
        bt.nup = r0.nup;
        bt.ndown = r0.ndown;
        bt.n = space_dim(r0);
        bt.da = binom_host(r0.L, r0.nup);
        bt.db = r0.kind == CF_DOCI ? bt.da : binom_host(r0.L, r0.ndown);
        bt.M = basis_size(r0, bt.n);
        bt.keep = std::max(1, std::min(4, bt.M - 1));
        bt.cap = capacity;
        bt.tol_f32 = dflt.tol_f32;
        bt.seed = dflt.seed;
        bt.maxit = 20000;
        bt.alloc();
        bt.add(first);
        bt.run();
    } catch (const cf::Error& e) {
        return cf::set_last_error(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return cf::set_last_error(CF_ERR_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return cf::set_last_error(CF_ERR_HIP, e.what());
    }
    return CF_OK;
}

}  // namespace cf

extern "C" int cf_solve_batch(cf_ctx* ctx, int count, const cf_request* reqs, cf_result* results) {
    try {
        CF_REQUIRE(ctx && (count == 0 || (reqs && results)), "NULL argument");
        CF_REQUIRE(count >= 0, "negative count");
        // CF_PREC_AUTO: FP64 (no operator of this batch stores matrix values)
        std::vector<cf_request> resolved(reqs, reqs + count);
        for (cf_request& q : resolved)
            if (q.precision == CF_PREC_AUTO) q.precision = CF_F64;
        reqs = resolved.data();
        // requests by shape, in request order
        std::map<uint64_t, std::vector<int>> groups;
        std::vector<uint64_t> order;
        for (int i = 0; i < count; ++i) {
            const cf_request& r = reqs[i];
            cf_result& o = results[i];
            std::memset(&o, 0, sizeof(o));
            o.tag = r.tag;
            o.status = CF_ERR_ARG;
            std::string why;
            if (!bd_accepts(r, &why)) {
                std::snprintf(o.error, sizeof(o.error), "%s", why.c_str());
                continue;
            }
            o.n = space_dim(r);
            const uint64_t key = bd_class_key(r);
            if (!groups.count(key)) order.push_back(key);
            groups[key].push_back(i);
        }
        BatchFeed feed;
        feed.done = [results](int id, const cf_result& r) { results[id] = r; };
        for (uint64_t key : order) {
            const std::vector<int>& idx = groups[key];
            for (size_t first = 0; first < idx.size(); first += kBdMaxLanes) {
                const size_t last = std::min(idx.size(), first + (size_t)kBdMaxLanes);
                std::vector<std::pair<int, cf_request>> in;
                for (size_t j = first; j < last; ++j) in.emplace_back(idx[j], reqs[idx[j]]);
                const int st = solve_batch_feed(ctx, (int)in.size(), in, feed);
                if (st != CF_OK) return st;
            }
        }
    } catch (const cf::Error& e) {
        return cf::set_last_error(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return cf::set_last_error(CF_ERR_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return cf::set_last_error(CF_ERR_HIP, e.what());
    }
    return CF_OK;
}
