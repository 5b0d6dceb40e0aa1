// C ABI of the ciflow_amd library (declarations and reference citations in
// include/ciflow_amd.h).
#include "../../include/ciflow_amd.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <fstream>
#include <limits>
#include <sstream>
#include <memory>
#include <string>
#include <algorithm>
#include <vector>

#include "capi_types.h"
#include "kernels.h"
#include "host.h"
#include "lanczos.h"
#include "rdm.h"
#include "sigma.h"
#include "batch.h"
#include "spmv_doci.h"
#include "spmv_ell.h"
#include "spmv_fci.h"

#include <rocblas/rocblas.h>
#include "transform.h"

namespace cf {
void gen_integrals(int L, uint64_t seed, double* h, double* V);
void noalloc_setup(int device);  // noalloc.hip
}

using namespace cf;

namespace {
int fail(int code, const std::string& msg) { return set_last_error(code, msg); }
}  // namespace

namespace {
// Copies ordered on the handle's stream (which does not synchronise with the
// null stream); return once the copy has completed.
void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
    if (!bytes) return;
    CF_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
    CF_HIP(hipStreamSynchronize(s));
}
}  // namespace

#define API_BEGIN try {
#define API_END                                              \
    }                                                        \
    catch (const cf::Error& e) {                             \
        return fail(e.code, e.what());                       \
    }                                                        \
    catch (const std::bad_alloc&) {                          \
        return fail(CF_ERR_NOMEM, "host allocation failed"); \
    }                                                        \
    catch (const std::exception& e) {                        \
        return fail(CF_ERR_HIP, e.what());                   \
    }                                                        \
    return CF_OK;

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
namespace cf {
hipStream_t ctx_stream(const cf_ctx* c) { return c->stream; }
int ctx_device(const cf_ctx* c) { return c->device; }
}  // namespace cf

namespace {
// Records HIP events around one kernel launch on the context stream.
struct KTimer {
    cf_ctx* c;
    const char* fam;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(cf_ctx* ctx, const char* family) : c(ctx), fam(family) {
        if (c->timing) {
            a = c->event();
            b = c->event();
            CF_HIP(hipEventRecord(a, c->stream));
        }
    }
    ~KTimer() {
        if (a) {
            (void)hipEventRecord(b, c->stream);
            c->timers[fam].pending.emplace_back(a, b);
        }
    }
};
}  // namespace

// One SpMV launch configuration: kernel 1 (row gather, `lanes` per row),
// 3 (vectorised row gather, `lanes` per row, `unroll`), 4 (FCI ELL, `ell`),
// 5 (DOCI hops, `doci`), 6 (DOCI blocks) or 7 (sliced rows, `lanes` = waves
// per 64-row chunk, `unroll`).
struct SpmvChoice {
    int kernel = 1;
    int lanes = 0;
    int unroll = 0;
    EllShape ell;
    DociShape doci;
};

struct cf_crs {
    cf_space* sp = nullptr;
    uint64_t n = 0, nnz = 0;
    int64_t nnz_upper = -1;
    double zero = 0.0;  // build threshold: elements with |v| <= zero are not stored
    bool want32 = false;  // FP32 value copy requested at build
    // Hamiltonian version of the space the values belong to; a newer space
    // version refills the derived copies and marks the CRS arrays stale (they
    // are rebuilt, pattern included, when something reads them)
    uint64_t version = 0;
    bool crs_stale = false;
    DBuf<int64_t> rowptr;
    DBuf<uint32_t> col;
    DBuf<double> val;
    DBuf<float> val32;
    int lpr = 64;     // lanes per row of the scalar row gather
    int kernel = 0;   // 0 auto, 1 row gather, 2 FCI staged, 3 vectorised row gather, 4 FCI ELL, 5 DOCI hops,
                      // 6 DOCI blocks, 7 sliced rows
    // automatic choice per value precision (row-length heuristic, or the autotuner's pick)
    SpmvChoice auto64, auto32;
    // FCI position-major complete rows (kernel 4), built on demand
    FciRowsArgs fargs{};
    bool has_fargs = false;
    DBuf<double> ell;
    DBuf<float> ell32;
    DBuf<uint16_t> jsub, jfull;
    DBuf<uint2> ent;
    EllArgs eargs;
    bool ell_ready = false;
    // DOCI position-major hop columns + pair tables (kernel 5)
    DBuf<uint32_t> dcol;
    DBuf<double> dvk;
    DBuf<float> dvk32;
    DociEllArgs dargs;
    bool doci_ready = false;
    // DOCI block tables (kernel 6)
    DBuf<DociBlock> dblk;
    DBuf<uint16_t> drank_lo, dlo_of;
    DBuf<uint32_t> dlo_off;
    DociBlockArgs bargs;
    bool blk_ready = false;
    // sliced copy of the CRS rows (kernel 7): 64-row chunks, quad offsets per
    // chunk; rebuilt from the CRS arrays whenever those are rebuilt
    DBuf<int64_t> sqoff;
    DBuf<uint32_t> scol;
    DBuf<double> sval;
    DBuf<float> sval32;
    DBuf<double> spart;  // range sums of a split launch (kSellMaxSplit x n, with fewer than 2^21 rows)
    bool sell_ready = false;
    int sell_waves = 16;  // waves per chunk when kernel 7 is set explicitly
    std::unique_ptr<FciSpmvPlan> plan;
};

namespace cf {
// One-string diagonal parts and excitation lists (values included) of a
// string table for the Hamiltonian H; the strings themselves do not depend on H.
void fill_strtab(cf_ctx* ctx, const cf_ham* H, StrTab* t) {
    launch_diag1(t->nel, t->s.p, t->d, H->dev(), t->d1.p, ctx->stream);
    if (t->E)
        launch_exc_lists(H->L, t->nel, t->s.p, t->d, ctx->binom.p, H->dev(), t->d1.p, t->E, t->S, t->full.p, t->sub.p,
                         ctx->stream);
}

// Diagonal of the space for its current Hamiltonian.
void fill_space_diag(cf_space* sp) {
    const cf_ham* ham = sp->ham;
    cf_ctx* c = ham->ctx;
    if (sp->kind == CF_DOCI)
        launch_doci_diag(sp->n, sp->ta->s.p, sp->ta->d1.p, ham->dev(), sp->diag.p, c->stream);
    else if (sp->kind == CF_FCI)
        launch_fci_diag(sp->da, sp->db, sp->ta->s.p, sp->tb->s.p, sp->ta->d1.p, sp->tb->d1.p, ham->dev(),
                        sp->diag.p, c->stream);
}

StrTab* make_strtab(cf_ctx* ctx, const cf_ham* H, int nel, bool lists) {
    std::unique_ptr<StrTab> t(new StrTab());
    t->L = H->L;
    t->nel = nel;
    t->d = binom_host(H->L, nel);
    CF_REQUIRE(t->d > 0 && t->d < (1ull << 32), "string space dimension must be in [1, 2^32)");
    t->s.alloc(t->d);
    t->d1.alloc(t->d);
    launch_unrank(H->L, nel, 0, t->d, ctx->binom.p, t->s.p, ctx->stream);
    if (lists) {
        const int nv = H->L - nel;
        const uint64_t ns = (uint64_t)nel * nv;
        const uint64_t nd = binom_host(nel, 2) * binom_host(nv, 2);
        CF_REQUIRE(1 + ns + nd <= 2048, "excitation list longer than 2048 entries");
        t->E = (int)(1 + ns + nd);
        t->S = (int)(1 + ns);
        t->full.alloc(t->d * t->E);
        t->sub.alloc(t->d * t->S);
    }
    fill_strtab(ctx, H, t.get());
    return t.release();
}

void make_doci_block_tables(cf_ctx* c, int L, int np, int w, cf_ctx::DociBlockTables& T) {
    const int nh = L - w;
    CF_REQUIRE(w >= 1 && w <= kDociMaxW && nh >= 1 && nh <= 20,
               "kernel 6: 1 <= w <= 12 low orbitals and 1..20 high orbitals");
    std::vector<DociBlock> blocks;
    for (uint32_t hi = 0; hi < (1u << nh); ++hi) {
        const int ch = __builtin_popcount(hi), cl = np - ch;
        if (cl < 0 || cl > w) continue;
        DociBlock b;
        b.hi = hi;
        b.c = (uint32_t)cl;
        b.R = (uint32_t)binom_host(w, cl);
        uint64_t base = 0;
        int mm = cl + 1;
        for (uint32_t t = hi; t; t &= t - 1) base += binom_host(w + __builtin_ctz(t), mm++);
        b.base = (uint32_t)base;
        blocks.push_back(b);
    }
    std::sort(blocks.begin(), blocks.end(), [](const DociBlock& x, const DociBlock& y) { return x.base < y.base; });
    const uint32_t W = 1u << w;
    std::vector<uint16_t> rank_lo(W), lo_of;
    std::vector<uint32_t> lo_off(w + 2, 0);
    for (int cl = 0; cl <= w; ++cl) {
        lo_off[cl] = (uint32_t)lo_of.size();
        for (uint32_t p = 0; p < W; ++p)  // increasing integers = rank order within a popcount
            if (__builtin_popcount(p) == cl) {
                rank_lo[p] = (uint16_t)(lo_of.size() - lo_off[cl]);
                lo_of.push_back((uint16_t)p);
            }
    }
    lo_off[w + 1] = (uint32_t)lo_of.size();
    T.blocks.alloc(blocks.size());
    T.rank_lo.alloc(W);
    T.lo_of.alloc(W);
    T.lo_off.alloc(w + 2);
    copy_sync(T.blocks.p, blocks.data(), blocks.size() * sizeof(DociBlock), hipMemcpyHostToDevice, c->stream);
    copy_sync(T.rank_lo.p, rank_lo.data(), W * sizeof(uint16_t), hipMemcpyHostToDevice, c->stream);
    copy_sync(T.lo_of.p, lo_of.data(), W * sizeof(uint16_t), hipMemcpyHostToDevice, c->stream);
    copy_sync(T.lo_off.p, lo_off.data(), (w + 2) * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream);
    DociBlockArgs& b = T.args;
    b.n = binom_host(L, np);
    b.L = L;
    b.np = np;
    b.w = w;
    b.blocks = T.blocks.p;
    b.nblocks = (uint32_t)blocks.size();
    b.rank_lo = T.rank_lo.p;
    b.lo_of = T.lo_of.p;
    b.lo_off = T.lo_off.p;
    b.binom = c->binom.p;
    b.diag = nullptr;
    b.zero = 0.0;
}
}  // namespace cf

namespace {
void ensure_staging(cf_ctx* c, uint64_t n) {
    if (c->hx.n < n) c->hx.alloc(n);
    if (c->hy.n < n) c->hy.alloc(n);
}

void sync_ham(cf_crs* m);
void ensure_crs(cf_crs* m);

// The configuration spmv() runs for a value precision: the explicit kernel if
// one was set, else the automatic choice.
SpmvChoice resolve_spmv(const cf_crs* m, int prec) {
    const SpmvChoice& a = prec == CF_F32 ? m->auto32 : m->auto64;
    SpmvChoice r = a;
    if (m->kernel) r.kernel = m->kernel;
    if (r.kernel == 2 && !m->plan) r.kernel = 1;
    if (r.kernel == 4 && !m->ell_ready) r.kernel = 1;
    if (r.kernel == 5 && !m->doci_ready) r.kernel = 1;
    if (r.kernel == 6 && !m->blk_ready) r.kernel = 1;
    if (r.kernel == 1) r.lanes = a.kernel == 1 && a.lanes && !m->kernel ? a.lanes : m->lpr;
    if (r.kernel == 3 && a.kernel != 3) {
        r.lanes = 16;
        r.unroll = 2;
    }
    if (r.kernel == 7 && a.kernel != 7) {
        r.lanes = m->sell_waves;
        r.unroll = 2;
    }
    return r;
}

// Sliced copy of the CRS rows for the current Hamiltonian (kernel 7): quads
// of each 64-row chunk's longest row, their exclusive scan, then the fill.
// With `optional`, a copy larger than half of the free device memory is not
// made (returns false).
bool make_sell(cf_crs* m, bool optional = false) {
    ensure_crs(m);
    if (m->sell_ready) return true;
    cf_ctx* c = m->sp->ham->ctx;
    const uint64_t nchunk = (m->n + 63) / 64;
    DBuf<uint32_t> quads(nchunk);
    if (!m->sqoff.p) m->sqoff.alloc(nchunk + 1);
    launch_sell_quads(m->n, m->rowptr.p, quads.p, c->stream);
    exclusive_scan_u32_to_i64(quads.p, m->sqoff.p, nchunk, c->stream);
    int64_t total = 0;
    CF_HIP(hipMemcpyAsync(&total, m->sqoff.p + nchunk, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipStreamSynchronize(c->stream));
    const uint64_t entries = (uint64_t)total * 256;
    if (optional && m->scol.n != entries) {
        size_t free_b = 0, total_b = 0;
        CF_HIP(hipMemGetInfo(&free_b, &total_b));
        if (entries * (12 + (m->want32 ? 4 : 0)) > free_b / 2) return false;
    }
    if (m->scol.n != entries) {
        m->scol.alloc(entries);
        m->sval.alloc(entries);
        if (m->want32) m->sval32.alloc(entries);
    }
    if (!m->spart.p && m->n < (1ull << 21)) m->spart.alloc(kSellMaxSplit * m->n);
    KTimer t(c, "crs_fill");
    launch_sell_fill(m->n, m->rowptr.p, m->col.p, m->val.p, m->want32 ? m->val32.p : nullptr, m->sqoff.p, m->scol.p,
                     m->sval.p, m->want32 ? m->sval32.p : nullptr, c->stream);
    CF_HIP(hipStreamSynchronize(c->stream));
    m->sell_ready = true;
    return true;
}

void spmv_with(cf_crs* m, int prec, const SpmvChoice& ch, const double* dx, double* dy) {
    cf_ctx* c = m->sp->ham->ctx;
    sync_ham(m);
    if (m->crs_stale && ch.kernel < 4) ensure_crs(m);
    if (ch.kernel == 7 && (m->crs_stale || !m->sell_ready)) make_sell(m);
    if (prec == CF_F32) {
        CF_REQUIRE(m->want32, "matrix was built without FP32 values");
        KTimer t(c, "spmv_f32");
        switch (ch.kernel) {
            case 5: launch_spmv_doci_f32(m->dargs, ch.doci, m->dvk32.p, dx, dy, c->stream); break;
            case 6: launch_spmv_doci_blk_f32(m->bargs, m->dvk32.p, dx, dy, c->stream); break;
            case 4: launch_spmv_ell_f32(m->eargs, ch.ell, m->ell32.p, dx, dy, c->stream); break;
            case 2: launch_spmv_fci_f32(m->plan->a, m->val32.p, dx, dy, c->stream); break;
            case 3:
                launch_spmv_v4_f32(m->n, m->rowptr.p, m->col.p, m->val32.p, dx, dy, ch.lanes, ch.unroll, c->stream);
                break;
            case 7:
                launch_spmv_sell_f32(m->n, m->sqoff.p, m->scol.p, m->sval32.p, dx, dy, m->spart.p, m->spart.n, ch.lanes,
                                     ch.unroll, c->stream);
                break;
            default: launch_spmv_f32(m->n, m->rowptr.p, m->col.p, m->val32.p, dx, dy, ch.lanes, c->stream);
        }
    } else {
        KTimer t(c, "spmv_f64");
        switch (ch.kernel) {
            case 5: launch_spmv_doci_f64(m->dargs, ch.doci, m->dvk.p, dx, dy, c->stream); break;
            case 6: launch_spmv_doci_blk_f64(m->bargs, m->dvk.p, dx, dy, c->stream); break;
            case 4: launch_spmv_ell_f64(m->eargs, ch.ell, m->ell.p, dx, dy, c->stream); break;
            case 2: launch_spmv_fci_f64(m->plan->a, m->val.p, dx, dy, c->stream); break;
            case 3:
                launch_spmv_v4_f64(m->n, m->rowptr.p, m->col.p, m->val.p, dx, dy, ch.lanes, ch.unroll, c->stream);
                break;
            case 7:
                launch_spmv_sell_f64(m->n, m->sqoff.p, m->scol.p, m->sval.p, dx, dy, m->spart.p, m->spart.n, ch.lanes,
                                     ch.unroll, c->stream);
                break;
            default: launch_spmv_f64(m->n, m->rowptr.p, m->col.p, m->val.p, dx, dy, ch.lanes, c->stream);
        }
    }
}

void spmv(cf_crs* m, int prec, const double* dx, double* dy) { spmv_with(m, prec, resolve_spmv(m, prec), dx, dy); }

// Position-major DOCI hop columns + pair tables (kernel 5).
// DOCI pair table (kernels 5 and 6): L x L elements, FP64 and FP32.
void make_doci_table(cf_crs* m) {
    if (m->dvk.p) return;
    cf_space* sp = m->sp;
    cf_ham* H = sp->ham;
    CF_REQUIRE(sp->kind == CF_DOCI, "the DOCI pair-hop SpMV needs a DOCI matrix");
    DociEllArgs& a = m->dargs;
    a.n = m->n;
    a.L = H->L;
    a.np = H->nup;
    a.nv = H->L - H->nup;
    a.s = sp->ta->s.p;
    a.diag = sp->diag.p;
    a.zero = m->zero;
    m->dvk.alloc((uint64_t)a.L * a.L);
    m->dvk32.alloc((uint64_t)a.L * a.L);
    launch_doci_pair_table(a, H->dev(), m->dvk.p, m->dvk32.p, H->ctx->stream);
}

void make_doci_ell(cf_crs* m) {
    make_doci_table(m);
    cf_ctx* c = m->sp->ham->ctx;
    DociEllArgs& a = m->dargs;
    m->dcol.alloc(std::max<uint64_t>(1, m->n * (uint64_t)a.np * (uint64_t)a.nv));
    a.col = m->dcol.p;
    {
        KTimer t(c, "doci_ell_fill");
        launch_doci_ell_fill(a, c->binom.p, m->dcol.p, c->stream);
    }
    m->doci_ready = true;
}

// Block tables of the block-staged DOCI SpMV (kernel 6) for w low orbitals.
void make_doci_blocks(cf_crs* m, int w) {
    make_doci_table(m);
    cf_space* sp = m->sp;
    cf_ctx::DociBlockTables T;
    make_doci_block_tables(sp->ham->ctx, sp->ham->L, sp->ham->nup, w, T);
    m->dblk = std::move(T.blocks);
    m->drank_lo = std::move(T.rank_lo);
    m->dlo_of = std::move(T.lo_of);
    m->dlo_off = std::move(T.lo_off);
    m->bargs = T.args;
    m->bargs.n = m->n;
    m->bargs.diag = sp->diag.p;
    m->bargs.zero = m->zero;
    m->blk_ready = true;
}

// Position-major complete FCI rows + beta index tables (kernel 4).
void make_ell(cf_crs* m) {
    cf_space* sp = m->sp;
    cf_ctx* c = sp->ham->ctx;
    CF_REQUIRE(sp->kind == CF_FCI && m->has_fargs, "the ELL SpMV needs an FCI matrix");
    CF_REQUIRE(sp->db <= kEllMaxDb, "the ELL SpMV needs d_beta <= 8192");
    const int Ea = sp->ta->E, Sb = sp->tb->S, Eb = sp->tb->E;
    const int nsa = sp->ta->S - 1, nda = Ea - 1 - nsa;
    const uint64_t R = (uint64_t)Eb + (uint64_t)nsa * Sb + nda;
    m->jsub.alloc((uint64_t)Sb * sp->db);
    m->jfull.alloc((uint64_t)Eb * sp->db);
    launch_ell_index(sp->db, Sb, Eb, sp->tb->sub.p, sp->tb->full.p, m->jsub.p, m->jfull.p, c->stream);
    m->ent.alloc(sp->da * (uint64_t)Ea);
    launch_ell_entries(sp->da, Ea, nda, nsa, sp->ta->full.p, sp->segoff.p, m->ent.p, c->stream);
    m->ell.alloc(m->n * R);
    if (m->want32) m->ell32.alloc(m->n * R);
    {
        KTimer t(c, "crs_fill");
        launch_fci_ell(m->fargs, m->ell.p, m->ell32.p, c->stream);
    }
    EllArgs& e = m->eargs;
    e.da = sp->da;
    e.db = sp->db;
    e.Ea = Ea;
    e.Sb = Sb;
    e.Eb = Eb;
    e.fa = sp->ta->full.p;
    e.segoff = sp->segoff.p;
    e.jsub = m->jsub.p;
    e.jfull = m->jfull.p;
    e.ent = m->ent.p;
    e.nda = nda;
    e.nsa = nsa;
    e.R = R;
    m->ell_ready = true;
}

// Staged FCI SpMV plan (alpha-block structure of the rows).
void make_fci_plan(cf_crs* m, int W) {
    cf_space* sp = m->sp;
    const int Ea = sp->ta->E, Sb = sp->tb->S, Eb = sp->tb->E;
    const int nsa = sp->ta->S - 1, nda = Ea - 1 - nsa;
    const int R = Eb + nsa * Sb + nda;
    std::unique_ptr<FciSpmvPlan> p(new FciSpmvPlan());
    p->init(sp->da, sp->db, Ea, sp->ta->full.p, sp->segoff.p, R, nda, Sb, W, sp->ham->ctx->stream);
    p->a.rowptr = m->rowptr.p;
    p->a.col = m->col.p;
    m->plan = std::move(p);
}

// Average row length decides how many lanes share a row (measured on MI355X:
// long FCI rows run best with 4 rows per wavefront, whose alpha-double
// gathers then coalesce; DOCI rows of ~150 entries with 2 rows per wavefront).
int pick_lanes_per_row(uint64_t n, uint64_t nnz) {
    double avg = n ? (double)nnz / (double)n : 0.0;
    if (avg >= 512) return 16;
    if (avg >= 128) return 32;
    if (avg >= 24) return 16;
    return 8;
}

// Count -> scan -> fill of the CRS arrays for the space's current
// Hamiltonian (CIMethod::build_parallel, CIMethod.cpp:489-561); reallocates
// col/val when the count changes.
void build_crs_arrays(cf_crs* m) {
    cf_space* sp = m->sp;
    cf_ctx* c = sp->ham->ctx;
    const double zero = m->zero;
    DBuf<uint32_t> cnt(m->n);
    if (!m->rowptr.p) m->rowptr.alloc(m->n + 1);
    const cf_ham* H = sp->ham;
    FciRowsArgs fa{};
    DociRowsArgs dargs{};
    DetRowsArgs targs{};
    // determinant lists with L <= 32 find each row's columns by excitation
    // lookup in the sorted (up, down) keys instead of n^2 pair tests
    const bool lookup = sp->kind == CF_DETS && H->L <= 32;
    DBuf<uint64_t> skey;
    DBuf<uint32_t> sidx;
    DetLookupArgs largs{};
    if (sp->kind == CF_DETS) {
        targs.n = sp->n;
        targs.nup = H->nup;
        targs.ndown = H->ndown;
        targs.up = sp->dup.p;
        targs.down = sp->ddown.p;
        targs.H = H->dev();
        targs.zero = zero;
        targs.rowcnt = cnt.p;
        if (lookup) {
            skey.alloc(sp->n);
            sidx.alloc(sp->n);
            sort_det_keys(sp->n, sp->dup.p, sp->ddown.p, skey.p, sidx.p, 32 + H->L, c->stream);
            largs.r = targs;
            largs.skey = skey.p;
            largs.sidx = sidx.p;
            largs.L = H->L;
            KTimer t(c, "crs_count");
            launch_det_rows_lookup(largs, false, c->stream);
        } else {
            KTimer t(c, "crs_count");
            launch_det_rows(targs, false, c->stream);
        }
    } else if (sp->kind == CF_FCI) {
        fa.da = sp->da;
        fa.db = sp->db;
        fa.Ea = sp->ta->E;
        fa.Sa = sp->ta->S;
        fa.Eb = sp->tb->E;
        fa.Sb = sp->tb->S;
        fa.sa = sp->ta->s.p;
        fa.sb = sp->tb->s.p;
        fa.fa = sp->ta->full.p;
        fa.fb = sp->tb->full.p;
        fa.subb = sp->tb->sub.p;
        fa.segoff = sp->segoff.p;
        fa.diag = sp->diag.p;
        fa.H = H->dev();
        fa.zero = zero;
        fa.row0 = 0;
        fa.nrows = m->n;
        fa.rowcnt = cnt.p;
        m->fargs = fa;
        m->fargs.rowcnt = nullptr;
        m->has_fargs = true;
        KTimer t(c, "crs_count");
        launch_fci_rows(fa, false, c->stream);
    } else {
        dargs.d = sp->n;
        dargs.L = H->L;
        dargs.np = H->nup;
        dargs.s = sp->ta->s.p;
        dargs.binom = c->binom.p;
        dargs.diag = sp->diag.p;
        dargs.H = H->dev();
        dargs.zero = zero;
        dargs.row0 = 0;
        dargs.nrows = m->n;
        dargs.rowcnt = cnt.p;
        KTimer t(c, "crs_count");
        launch_doci_rows(dargs, false, c->stream);
    }
    exclusive_scan_u32_to_i64(cnt.p, m->rowptr.p, m->n, c->stream);
    int64_t nnz = 0;
    copy_sync(&nnz, m->rowptr.p + m->n, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
    m->nnz = (uint64_t)nnz;
    m->col.alloc(m->nnz);
    m->val.alloc(m->nnz);
    if (m->want32) m->val32.alloc(m->nnz);
    if (sp->kind == CF_DETS) {
        targs.rowptr = m->rowptr.p;
        targs.col = m->col.p;
        targs.val = m->val.p;
        targs.val32 = m->val32.p;
        KTimer t(c, "crs_fill");
        if (lookup && m->nnz <= 0x7fffffffull) {
            largs.r.rowcnt = nullptr;
            largs.r.rowptr = m->rowptr.p;
            largs.r.col = m->col.p;
            largs.r.val = m->val.p;
            launch_det_rows_lookup(largs, true, c->stream);
            sort_rows_by_column(m->n, m->nnz, m->rowptr.p, m->col.p, m->val.p, c->stream);
            if (m->want32) launch_to_f32(m->nnz, m->val.p, m->val32.p, c->stream);
        } else {
            launch_det_rows(targs, true, c->stream);
        }
    } else if (sp->kind == CF_FCI) {
        fa.rowptr = m->rowptr.p;
        fa.col = m->col.p;
        fa.val = m->val.p;
        fa.val32 = m->val32.p;
        KTimer t(c, "crs_fill");
        launch_fci_rows(fa, true, c->stream);
    } else {
        dargs.rowptr = m->rowptr.p;
        dargs.col = m->col.p;
        dargs.val = m->val.p;
        dargs.val32 = m->val32.p;
        KTimer t(c, "crs_fill");
        launch_doci_rows(dargs, true, c->stream);
    }
    CF_HIP(hipStreamSynchronize(c->stream));
    m->lpr = pick_lanes_per_row(m->n, m->nnz);
    m->nnz_upper = -1;
    m->crs_stale = false;
}

// Brings a matrix up to its space's Hamiltonian after cf_space_set_ham: the
// pair table (kernel 5) and the position-major values (kernel 4) are refilled
// in place -- their layouts do not depend on the integrals -- and the CRS
// arrays are marked stale, to be rebuilt (pattern included, since elements
// can cross the drop threshold) when something reads them.
void sync_ham(cf_crs* m) {
    cf_space* sp = m->sp;
    if (m->version == sp->version) return;
    cf_ctx* c = sp->ham->ctx;
    const HamDev H = sp->ham->dev();
    if (m->dvk.p) launch_doci_pair_table(m->dargs, H, m->dvk.p, m->dvk32.p, c->stream);
    if (m->has_fargs) m->fargs.H = H;
    if (m->ell_ready) {
        KTimer t(c, "crs_fill");
        launch_fci_ell(m->fargs, m->ell.p, m->ell32.p, c->stream);
    }
    m->plan.reset();  // the staged kernel reads the CRS arrays; rebuilt on request
    m->sell_ready = false;  // the sliced copy too, after the CRS arrays
    m->crs_stale = true;
    m->version = sp->version;
}

// The CRS arrays for the current Hamiltonian.
void ensure_crs(cf_crs* m) {
    sync_ham(m);
    if (m->crs_stale) {
        build_crs_arrays(m);
        m->lpr = pick_lanes_per_row(m->n, m->nnz);
    }
}

// ---------------------------------------------------------------------------
// Krylov backends
// ---------------------------------------------------------------------------
struct DeviceBackend : KrylovBackend {
    cf_ctx* c;
    uint64_t n;
    int M;
    DBuf<double> Vb, w, partial, cvec, c2, Yd, scal, dalpha, dbeta, pnorm;
    bool first_ = true;  // next orthogonalisation opens a cycle
    int nb;
    bool fused;  // step kernels reduce the partials themselves (n <= kFusedMaxN)
    int G;       // workgroups of the wide product in the fused step
    std::function<void(const double*, double*)> op;
    double* hbuf;  // the context's pinned read-back buffer

    DeviceBackend(cf_ctx* ctx, uint64_t n_, int ncv) : c(ctx), n(n_) {
        M = (uint64_t)ncv > n ? (int)n : ncv;
        Vb.alloc((size_t)(M + 1) * n);
        w.alloc(n);
        nb = multi_dot_blocks(n);
        fused = n <= kFusedMaxN;
        G = fused_dot_blocks(n);
        partial.alloc((size_t)64 * std::max(nb, G));
        pnorm.alloc(256);

        cvec.alloc(64);
        c2.alloc(64);
        Yd.alloc(64 * 64);
        scal.alloc(4);
        dalpha.alloc(64);
        dbeta.alloc(64);
        hbuf = c->pinned;
    }
    double* v(int i) { return Vb.p + (size_t)i * n; }
    // Partials + a separate fixed-order reduction: a last-block reduction inside
    // one launch needs a device-scope release fence per block, which on this GPU
    // writes back the XCD's L2 and made the fused form 2.5-10x slower.
    void dots(int k, const double* base, const double* x, double* out) {
        launch_multi_dot(n, k, base, x, partial.p, nb, c->stream);
        launch_reduce_partials(k, nb, partial.p, out, c->stream);
    }
    double fetch(const double* d) {
        CF_HIP(hipMemcpyAsync(hbuf, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CF_HIP(hipStreamSynchronize(c->stream));
        return hbuf[0];
    }
    bool deferred() const override { return true; }
    // The first step of a cycle (w couples to every kept Ritz vector) takes two
    // classical Gram-Schmidt passes; later steps remove the three-term part
    // alpha_j v_j + beta_{j-1} v_{j-1} first and then take one full pass, which
    // halves the basis traffic.  alpha_j = (v_j . w) + correction of the pass.
    void orth_deferred(int j) override {
        if (fused && j + 1 <= 64) {
            // five launches: product, three-term update (or first pass), product,
            // update + |w|^2 parts, normalisation; no separate reductions
            if (first_ || j == 0) {
                launch_multi_dot_wide(n, j + 1, Vb.p, w.p, partial.p, G, c->stream);
                launch_multi_axpy_r(n, j + 1, G, partial.p, Vb.p, w.p, cvec.p, pnorm.p, c->stream);
            } else {
                launch_multi_dot_wide(n, 1, v(j), w.p, partial.p, G, c->stream);
                launch_local3_r(n, G, partial.p, v(j), dbeta.p + (j - 1), v(j - 1), w.p, cvec.p + j, c->stream);
            }
            first_ = false;
            launch_multi_dot_wide(n, j + 1, Vb.p, w.p, partial.p, G, c->stream);
            launch_multi_axpy_r(n, j + 1, G, partial.p, Vb.p, w.p, c2.p, pnorm.p, c->stream);
            launch_finish_step_r(n, j, cvec.p, c2.p, pnorm.p, w.p, v(j + 1), dalpha.p, dbeta.p, c->stream);
            return;
        }
        if (first_ || j == 0) {
            dots(j + 1, Vb.p, w.p, cvec.p);
            launch_multi_axpy(n, j + 1, Vb.p, cvec.p, w.p, c->stream);
        } else {
            dots(1, v(j), w.p, cvec.p + j);
            launch_local3(n, cvec.p + j, v(j), dbeta.p + (j - 1), v(j - 1), w.p, c->stream);
        }
        first_ = false;
        dots(j + 1, Vb.p, w.p, c2.p);
        launch_multi_axpy(n, j + 1, Vb.p, c2.p, w.p, c->stream);
        dots(1, w.p, w.p, scal.p);
        launch_finish_step(n, j, cvec.p, c2.p, scal.p, w.p, v(j + 1), dalpha.p, dbeta.p, c->stream);
    }
    void fetch_coeffs(int j0, int j1, double* alpha, double* beta) override {
        const int k = j1 - j0;
        CF_HIP(hipMemcpyAsync(hbuf, dalpha.p + j0, k * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CF_HIP(hipMemcpyAsync(hbuf + 64, dbeta.p + j0, k * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CF_HIP(hipStreamSynchronize(c->stream));
        for (int i = 0; i < k; ++i) {
            alpha[i] = hbuf[i];
            beta[i] = hbuf[64 + i];
        }
    }
    void start_random(uint64_t seed) override {
        launch_fill_start(n, seed, v(0), c->stream);
        normalize0();
    }
    void start_keep() override { normalize0(); }
    void apply(int j) override { op(v(j), w.p); }
    void orth(int j, double* alpha, double* beta) override {
        dots(j + 1, Vb.p, w.p, cvec.p);
        launch_multi_axpy(n, j + 1, Vb.p, cvec.p, w.p, c->stream);
        dots(j + 1, Vb.p, w.p, c2.p);
        launch_multi_axpy(n, j + 1, Vb.p, c2.p, w.p, c->stream);
        dots(1, w.p, w.p, scal.p);
        CF_HIP(hipMemcpyAsync(&hbuf[0], cvec.p + j, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CF_HIP(hipMemcpyAsync(&hbuf[1], c2.p + j, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CF_HIP(hipMemcpyAsync(&hbuf[2], scal.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CF_HIP(hipStreamSynchronize(c->stream));
        *alpha = hbuf[0] + hbuf[1];
        *beta = std::sqrt(std::max(0.0, hbuf[2]));
    }
    void set_next(int j1, double beta) override { launch_scale_copy(n, w.p, 1.0 / beta, v(j1), c->stream); }
    void rotate(int m, int k, const double* Y) override {
        // staged through page-locked memory (m <= 63); the wait keeps the staging area
        // valid until the copy has run
        double* ys = hbuf + 256;
        std::memcpy(ys, Y, sizeof(double) * m * m);
        CF_HIP(hipMemcpyAsync(Yd.p, ys, sizeof(double) * m * m, hipMemcpyHostToDevice, c->stream));
        launch_rotate_basis(n, m, k, Vb.p, Yd.p, c->stream);
        CF_HIP(hipStreamSynchronize(c->stream));
        first_ = true;
    }
    void move(int from, int to) override { launch_copy(n, v(from), v(to), c->stream); }
    void normalize0() override {
        first_ = true;
        dots(1, v(0), v(0), scal.p);
        double s = fetch(scal.p);
        launch_scale_copy(n, v(0), 1.0 / std::sqrt(s), v(0), c->stream);
    }
    void rayleigh(double* theta, double* resid) override {
        dots(1, v(0), w.p, scal.p);
        double th = fetch(scal.p);
        hbuf[128] = th;  // page-locked source; the fetch below waits for the copy
        CF_HIP(hipMemcpyAsync(cvec.p, hbuf + 128, sizeof(double), hipMemcpyHostToDevice, c->stream));
        launch_multi_axpy(n, 1, v(0), cvec.p, w.p, c->stream);  // w -= theta v0
        dots(1, w.p, w.p, scal.p);
        double r2 = fetch(scal.p);
        *theta = th;
        *resid = std::sqrt(std::max(0.0, r2));
    }
};

// The solve's vector becomes the space's resident eigenvector.  With a host
// buffer it is read back, given the sign convention (largest-magnitude
// component positive) and returned; without one it stays on the device as it
// is and the sign is fixed when it is read (cf_space_eigvec), which saves the
// served solves two copies of n doubles.
void keep_eigvec(cf_space* sp, uint64_t n, const double* dvec, double* eigvec) {
    cf_ctx* c = sp->ham->ctx;
    if (sp->vec.n != n) sp->vec.alloc(n);
    if (!eigvec) {
        launch_copy(n, dvec, sp->vec.p, c->stream);
        CF_HIP(hipStreamSynchronize(c->stream));
        sp->vec_valid = true;
        sp->vec_sign_pending = true;
        return;
    }
    std::vector<double> hv(n);
    CF_HIP(hipMemcpyAsync(hv.data(), dvec, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipStreamSynchronize(c->stream));
    sign_fix(hv);
    copy_sync(sp->vec.p, hv.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    sp->vec_valid = true;
    sp->vec_sign_pending = false;
    std::memcpy(eigvec, hv.data(), n * sizeof(double));
}
}  // namespace

// ---------------------------------------------------------------------------
// API
// ---------------------------------------------------------------------------
namespace {
void check_shape(const cf_crs* m, const SpmvChoice& ch) {
    if (ch.kernel == 1) {
        CF_REQUIRE(ch.lanes == 8 || ch.lanes == 16 || ch.lanes == 32 || ch.lanes == 64,
                   "kernel 1: lanes per row must be 8, 16, 32 or 64");
    } else if (ch.kernel == 3) {
        CF_REQUIRE(ch.lanes == 4 || ch.lanes == 8 || ch.lanes == 16 || ch.lanes == 32,
                   "kernel 3: lanes per row must be 4, 8, 16 or 32");
        CF_REQUIRE(ch.unroll == 1 || ch.unroll == 2, "kernel 3: unroll must be 1 or 2");
    } else if (ch.kernel == 4) {
        CF_REQUIRE(m->ell_ready, "kernel 4 needs the FCI position-major copy (cf_crs_set_spmv_kernel(m, 4, 0))");
        const int t = ch.ell.threads, v = ch.ell.vec;
        CF_REQUIRE(v == 0 || v == 1 || v == 2 || v == 4, "kernel 4: strings per thread must be 0, 1, 2 or 4");
        if (v == 0) CF_REQUIRE(t == 256 || t == 512 || t == 1024, "kernel 4, vec 0: workgroup size 256, 512 or 1024");
        else CF_REQUIRE(t % 64 == 0 && t >= 64 && t <= 1024, "kernel 4: workgroup size must be a multiple of 64 <= 1024");
        CF_REQUIRE(v == 0 || m->sp->db % v == 0, "kernel 4: d_beta must be a multiple of the strings per thread");
        CF_REQUIRE(v == 0 || (m->sp->db + t - 1) / t <= 16, "kernel 4: workgroup too small for d_beta");
    } else if (ch.kernel == 5) {
        CF_REQUIRE(m->doci_ready, "kernel 5 needs a DOCI matrix");
        const int t = ch.doci.threads;
        CF_REQUIRE(t % 64 == 0 && t >= 64 && t <= 1024, "kernel 5: workgroup size must be a multiple of 64 <= 1024");
        const int u = ch.doci.sub;
        CF_REQUIRE(u == 1 || u == 2 || u == 4 || u == 8, "kernel 5: lanes per row must be 1, 2, 4 or 8");
    } else if (ch.kernel == 6) {
        CF_REQUIRE(m->blk_ready, "kernel 6 needs a DOCI matrix");
    } else if (ch.kernel == 7) {
        const int w = ch.lanes;
        CF_REQUIRE(w == 1 || w == 2 || w == 4 || w == 8 || w == 16, "kernel 7: waves per chunk must be 1, 2, 4, 8 or 16");
        CF_REQUIRE(ch.unroll == 1 || ch.unroll == 2 || (ch.unroll == 4 && w >= 4),
                   "kernel 7: unroll must be 1 or 2 (4 with 4, 8 or 16 waves)");
    } else {
        CF_REQUIRE(false, "shape kernel must be 1, 3, 4, 5, 6 or 7");
    }
}

SpmvChoice make_choice(int kernel, int p1, int p2, uint64_t db, int prec) {
    SpmvChoice ch;
    ch.kernel = kernel;
    if (kernel == 4) {
        ch.ell.threads = p1 > 0 ? p1 : ell_default_shape(db, prec == CF_F32 ? 4 : 8).threads;
        ch.ell.vec = p2;
    } else if (kernel == 5) {
        ch.doci.threads = p1 > 0 ? p1 : DociShape{}.threads;
        ch.doci.sub = p2;
    } else if (kernel == 7) {
        ch.lanes = p1 > 0 ? p1 : 16;
        ch.unroll = p2 > 0 ? p2 : 2;
    } else {
        ch.lanes = p1;
        ch.unroll = p2;
    }
    return ch;
}
}  // namespace

extern "C" {

int cf_ctx_create(int device, cf_ctx** out) {
    API_BEGIN
    CF_REQUIRE(out != nullptr, "out is NULL");
    int ndev = 0;
    CF_HIP(hipGetDeviceCount(&ndev));
    CF_REQUIRE(device >= 0 && device < ndev, "device index out of range");
    std::unique_ptr<cf_ctx> c(new cf_ctx());
    c->device = device;
    CF_HIP(hipSetDevice(device));
    // keep freed blocks in the device's stream-ordered pool for reuse
    hipMemPool_t mp = nullptr;
    CF_HIP(hipDeviceGetDefaultMemPool(&mp, device));
    uint64_t keep = UINT64_MAX;
    CF_HIP(hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &keep));
    CF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->use();
    CF_HIP(hipHostMalloc((void**)&c->pinned, cf_ctx::kPinned * sizeof(double), hipHostMallocDefault));
    std::vector<uint64_t> hb((kMaxL + 1) * kBinomCols, 0);
    for (int k = 0; k <= kMaxL; ++k)
        for (int m = 0; m < kBinomCols; ++m) hb[k * kBinomCols + m] = binom_host(k, m);
    c->binom.alloc(hb.size());
    copy_sync(c->binom.p, hb.data(), hb.size() * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream);
    noalloc_setup(device);
    *out = c.release();
    API_END
}

int cf_ctx_destroy(cf_ctx* ctx) {
    API_BEGIN
    if (ctx) {
        ctx->use();
        (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    }
    API_END
}

int cf_ctx_synchronize(cf_ctx* ctx) {
    API_BEGIN
    CF_REQUIRE(ctx, "ctx is NULL");
    ctx->use();
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_ctx_enable_timing(cf_ctx* ctx, int enable) {
    API_BEGIN
    CF_REQUIRE(ctx, "ctx is NULL");
    ctx->timing = enable != 0;
    API_END
}

int cf_ctx_kernel_time(cf_ctx* ctx, const char* family, int64_t* launches, double* total_ms) {
    API_BEGIN
    CF_REQUIRE(ctx && family, "NULL argument");
    ctx->use();
    auto it = ctx->timers.find(family);
    if (it == ctx->timers.end()) {
        if (launches) *launches = 0;
        if (total_ms) *total_ms = 0.0;
    } else {
        ctx->flush(it->second);
        if (launches) *launches = it->second.launches;
        if (total_ms) *total_ms = it->second.ms;
    }
    API_END
}

int cf_ctx_stream_copy(cf_ctx* ctx, uint64_t bytes, int reps) {
    API_BEGIN
    CF_REQUIRE(ctx && bytes >= 8 && reps >= 1, "bad arguments");
    ctx->use();
    const uint64_t n = bytes / 8;
    DBuf<double> a(n), b(n);
    CF_HIP(hipMemsetAsync(a.p, 0, n * 8, ctx->stream));
    for (int r = 0; r < reps; ++r) launch_copy(n, a.p, b.p, ctx->stream);
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_ctx_reset_timing(cf_ctx* ctx) {
    API_BEGIN
    CF_REQUIRE(ctx, "ctx is NULL");
    ctx->use();
    for (auto& kv : ctx->timers) {
        ctx->flush(kv.second);
        kv.second.launches = 0;
        kv.second.ms = 0.0;
    }
    API_END
}

int cf_ham_create(cf_ctx* ctx, int L, int nup, int ndown, const double* h, const double* V, double econst,
                  cf_ham** out) {
    API_BEGIN
    CF_REQUIRE(ctx && h && V && out, "NULL argument");
    CF_REQUIRE(L >= 1 && L <= kMaxL, "L must be in [1, 64]");
    CF_REQUIRE(nup >= 0 && nup <= L && ndown >= 0 && ndown <= L, "electron counts out of range");
    CF_REQUIRE(nup < kBinomCols && ndown < kBinomCols, "at most 32 electrons per spin");
    ctx->use();
    std::unique_ptr<cf_ham> H(new cf_ham());
    H->ctx = ctx;
    H->L = L;
    H->nup = nup;
    H->ndown = ndown;
    H->econst = econst;
    const size_t L2 = (size_t)L * L, L3 = L2 * L, L4 = L3 * L;
    std::vector<double> W(L3), Jm(L2);
    for (int p = 0; p < L; ++p)
        for (int q = 0; q < L; ++q)
            for (int d = 0; d < L; ++d) W[((size_t)p * L + q) * L + d] = V[(((size_t)p * L + d) * L + q) * L + d];
    for (int k = 0; k < L; ++k)
        for (int l = 0; l < L; ++l) Jm[(size_t)k * L + l] = V[(((size_t)k * L + l) * L + k) * L + l];
    H->dh.alloc(L2);
    H->dV.alloc(L4);
    H->dW.alloc(L3);
    H->dJm.alloc(L2);
    copy_sync(H->dh.p, h, L2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    copy_sync(H->dV.p, V, L4 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    copy_sync(H->dW.p, W.data(), L3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    copy_sync(H->dJm.p, Jm.data(), L2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    *out = H.release();
    API_END
}

int cf_ham_destroy(cf_ham* ham) {
    API_BEGIN
    if (ham) {
        ham->ctx->use();
        delete ham;
    }
    API_END
}

int cf_space_create(cf_ham* ham, int kind, cf_space** out) {
    API_BEGIN
    CF_REQUIRE(ham && out, "NULL argument");
    CF_REQUIRE(kind == CF_FCI || kind == CF_DOCI, "kind must be CF_FCI or CF_DOCI");
    cf_ctx* c = ham->ctx;
    c->use();
    std::unique_ptr<cf_space> sp(new cf_space());
    sp->ham = ham;
    sp->kind = kind;
    if (kind == CF_DOCI) {
        CF_REQUIRE(ham->nup == ham->ndown, "DOCI needs nup == ndown (pairs)");
        sp->ta.reset(make_strtab(c, ham, ham->nup, false));
        sp->tb = sp->ta.get();
        sp->da = sp->db = sp->ta->d;
        sp->n = sp->da;
        sp->diag.alloc(sp->n);
        fill_space_diag(sp.get());
    } else {
        sp->ta.reset(make_strtab(c, ham, ham->nup, true));
        if (ham->ndown == ham->nup) {
            sp->tb = sp->ta.get();
        } else {
            sp->tb_own.reset(make_strtab(c, ham, ham->ndown, true));
            sp->tb = sp->tb_own.get();
        }
        sp->da = sp->ta->d;
        sp->db = sp->tb->d;
        CF_REQUIRE(sp->da * sp->db < (1ull << 32), "FCI dimension must be < 2^32");
        sp->n = sp->da * sp->db;
        sp->segoff.alloc(sp->da * (uint64_t)(sp->ta->E + 1));
        launch_fci_segoff(sp->da, sp->ta->E, sp->ta->full.p, sp->tb->E, sp->tb->S, sp->segoff.p, c->stream);
        sp->diag.alloc(sp->n);
        fill_space_diag(sp.get());
    }
    CF_HIP(hipStreamSynchronize(c->stream));
    *out = sp.release();
    API_END
}

int cf_space_create_dets(cf_ham* ham, uint64_t n, const uint64_t* up, const uint64_t* down, cf_space** out) {
    API_BEGIN
    CF_REQUIRE(ham && out && (n == 0 || (up && down)), "NULL argument");
    CF_REQUIRE(n >= 1 && n < (1ull << 32), "determinant count must be in [1, 2^32)");
    const int L = ham->L;
    const uint64_t mask = L >= 64 ? ~0ull : ((1ull << L) - 1);
    uint64_t ascending = 0;
    for (uint64_t i = 0; i < n; ++i) {
        CF_REQUIRE((up[i] & ~mask) == 0 && (down[i] & ~mask) == 0,
                   "determinant " + std::to_string(i) + " occupies an orbital >= L");
        CF_REQUIRE(__builtin_popcountll(up[i]) == ham->nup && __builtin_popcountll(down[i]) == ham->ndown,
                   "determinant " + std::to_string(i) + " does not hold nup alpha and ndown beta electrons");
        if (i && (up[i] > up[i - 1] || (up[i] == up[i - 1] && down[i] > down[i - 1]))) ++ascending;
    }
    cf_ctx* c = ham->ctx;
    c->use();
    std::unique_ptr<cf_space> sp(new cf_space());
    sp->ham = ham;
    sp->kind = CF_DETS;
    sp->n = n;
    sp->da = n;
    sp->db = 1;
    sp->ascending = ascending;
    sp->dup.alloc(n);
    sp->ddown.alloc(n);
    copy_sync(sp->dup.p, up, n * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream);
    copy_sync(sp->ddown.p, down, n * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream);
    *out = sp.release();
    API_END
}

int cf_space_destroy(cf_space* sp) {
    API_BEGIN
    if (sp) {
        sp->ham->ctx->use();
        delete sp;
    }
    API_END
}

int cf_space_set_ham(cf_space* sp, cf_ham* ham) {
    API_BEGIN
    CF_REQUIRE(sp && ham, "NULL argument");
    CF_REQUIRE(ham->ctx == sp->ham->ctx, "the Hamiltonian belongs to another context");
    CF_REQUIRE(ham->L == sp->ham->L && ham->nup == sp->ham->nup && ham->ndown == sp->ham->ndown,
               "the Hamiltonian must have the space's L, nup and ndown");
    cf_ctx* c = ham->ctx;
    c->use();
    sp->ham = ham;
    if (sp->kind != CF_DETS) {  // determinant-list rows read the integrals directly
        fill_strtab(c, ham, sp->ta.get());
        if (sp->tb_own) fill_strtab(c, ham, sp->tb_own.get());
        fill_space_diag(sp);
    }
    sp->vec_valid = false;
    ++sp->version;
    CF_HIP(hipStreamSynchronize(c->stream));
    API_END
}

int cf_space_dims(const cf_space* sp, uint64_t* n, uint64_t* dim_up, uint64_t* dim_down) {
    API_BEGIN
    CF_REQUIRE(sp, "space is NULL");
    if (n) *n = sp->n;
    if (dim_up) *dim_up = sp->da;
    if (dim_down) *dim_down = sp->db;
    API_END
}

int cf_strings_unrank(cf_ctx* ctx, int L, int nel, uint64_t first, uint64_t count, uint64_t* out) {
    API_BEGIN
    CF_REQUIRE(ctx && out, "NULL argument");
    CF_REQUIRE(L >= 1 && L <= kMaxL && nel >= 0 && nel <= L && nel < kBinomCols, "bad L / nel");
    CF_REQUIRE(first + count <= binom_host(L, nel), "rank range beyond the space");
    ctx->use();
    DBuf<str_t> d(count);
    launch_unrank(L, nel, first, count, ctx->binom.p, d.p, ctx->stream);
    CF_HIP(hipMemcpyAsync(out, d.p, count * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_strings_rank(cf_ctx* ctx, int L, int nel, const uint64_t* strings, uint64_t count, uint32_t* ranks) {
    API_BEGIN
    CF_REQUIRE(ctx && strings && ranks, "NULL argument");
    CF_REQUIRE(L >= 1 && L <= kMaxL && nel >= 0 && nel < kBinomCols, "bad L / nel");
    const uint64_t full = L == 64 ? ~0ull : ((1ull << L) - 1);
    for (uint64_t i = 0; i < count; ++i)
        CF_REQUIRE((strings[i] & ~full) == 0 && __builtin_popcountll(strings[i]) == nel,
                   "string outside the n-of-L space");
    ctx->use();
    DBuf<str_t> d(count);
    DBuf<uint32_t> r(count);
    CF_HIP(hipMemcpyAsync(d.p, strings, count * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
    launch_rank(L, nel, d.p, count, ctx->binom.p, r.p, ctx->stream);
    CF_HIP(hipMemcpyAsync(ranks, r.p, count * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_crs_build(cf_space* sp, int value_prec, double zero, cf_crs** out) {
    API_BEGIN
    CF_REQUIRE(sp && out, "NULL argument");
    CF_REQUIRE(value_prec == CF_F64 || value_prec == CF_F32, "value_prec must be CF_F64 or CF_F32");
    CF_REQUIRE(zero >= 0.0, "zero threshold must be >= 0");
    cf_ctx* c = sp->ham->ctx;
    c->use();
    std::unique_ptr<cf_crs> m(new cf_crs());
    m->sp = sp;
    m->n = sp->n;
    m->zero = zero;
    CF_REQUIRE(m->n < (1ull << 32), "matrix dimension must be < 2^32 (u32 columns)");
    m->want32 = value_prec == CF_F32;
    m->version = sp->version;
    const cf_ham* H = sp->ham;
    // DOCI matrices that run the block-staged kernel (below) need no CRS for
    // products: its arrays (pattern and values) are built when something reads
    // them (export, counts, the CRS kernels), not here
    const int np = H->nup, nv = H->L - H->nup;
    const bool defer = sp->kind == CF_DOCI && m->n >= 60000 && H->L - doci_block_w(H->L) <= 20;
    const uint64_t doci_row = 1 + (uint64_t)np * (uint64_t)nv;
    if (defer) {
        m->crs_stale = true;
        m->lpr = pick_lanes_per_row(m->n, m->n * doci_row);
    } else {
        build_crs_arrays(m.get());
    }
    // measured on MI355X (FCI L=11/12, DOCI L=20/24): vectorised streams win for
    // FP32 values at every row length and for FP64 rows of >= 512 entries
    const double avg = defer ? (double)doci_row : m->n ? (double)m->nnz / (double)m->n : 0.0;
    m->auto64.kernel = m->auto32.kernel = 1;
    m->auto64.lanes = m->auto32.lanes = m->lpr;
    if (avg >= 512) m->auto64 = SpmvChoice{3, 16, 2, {}, {}};
    if (avg >= 128) m->auto32 = SpmvChoice{3, 8, 2, {}, {}};
    else if (avg >= 24) m->auto32 = SpmvChoice{3, 16, 1, {}, {}};
    // FCI matrices with enough alpha strings to fill the GPU also get the
    // position-major copy (kernel 4): measured 2.38 ms against 5.15 ms (FP64) and
    // 1.53 ms against 3.45 ms (FP32) at L = 12, 0.52 / 0.90 ms at L = 11; at
    // L = 10 (252 alpha strings) the row gather stays faster
    if (sp->kind == CF_FCI && sp->da >= 400 && sp->db <= kEllMaxDb) {
        make_ell(m.get());
        m->auto64 = SpmvChoice{4, 0, 0, ell_default_shape(sp->db, 8), {}};
        m->auto32 = SpmvChoice{4, 0, 0, ell_default_shape(sp->db, 4), {}};
    }
    // DOCI matrices: the pair elements come from an L x L table in LDS instead
    // of the 8/12-byte element stream of the row gather.  Measured on MI355X
    // (FP64 / FP32 values): L = 24 block-staged 0.262 / 0.262 ms, hop columns
    // 0.472 / 0.477, row gather 1.468 / 1.076; L = 22 68 / 69 against 121 / 125
    // (hop columns) and 312 / 226 us; L = 20 18.7 / 19.1 against 33.1 / 33.8
    // and 73 / 52 us; L = 19 (92,378 rows) 15.8 / 15.7 against 21.1 / 21.5 and
    // 31.7 / 23.7 us; L = 18 (48,620 rows) hop columns 13.2 against 14.5
    // (block-staged) and 15.7 / 11.5 us (row gather).  Below that the launch
    // dominates and the row gather is as fast, so smaller matrices build on
    // request.
    if (sp->kind == CF_DOCI && m->n >= 40000) {
        const int L = sp->ham->L, w = doci_block_w(L);
        if (m->n >= 60000 && L - w <= 20) {
            make_doci_blocks(m.get(), w);
            m->auto64 = m->auto32 = SpmvChoice{6, 0, 0, {}, {}};
        } else {
            make_doci_ell(m.get());
            const SpmvChoice k5{5, 0, 0, {}, doci_default_shape(m->n)};
            m->auto64 = k5;
        }
    }
    // Determinant lists in (up, down) order (at least 90% of neighbouring
    // pairs ascending), and FCI spaces left on the row gather (too few alpha
    // strings for the position-major copy) with >= 8,192 rows, get the sliced
    // copy (kernel 7) when it fits in half of the free device memory: their
    // neighbouring rows have neighbouring columns.  Measured on MI355X (FP64 /
    // FP32 values): CAS-like lists sorted by (up, down), 101,640 determinants
    // of L=12 0.242 / 0.146 ms against 0.361 / 0.253 ms (vectorised row
    // gather), 1,029,600 of L=16 4.76 / 2.93 ms against 6.81 / 4.57 ms; FCI
    // L=10 (63,504 rows) 0.120 / 0.077 ms against 0.165 / 0.122 ms, FCI L=9
    // (15,876) 26 / 16 us against 27 / 24 us.  In random order the same
    // CAS-like list runs 8-11% slower on the sliced kernel, a random 10^5 subset
    // of L=14 15-30%, so unsorted lists keep the row gather.  DOCI rows (65-82
    // entries at L=16-18) are too short to gain.
    const bool gather = m->auto64.kernel == 1 || m->auto64.kernel == 3;
    const bool sorted_list = sp->kind == CF_DETS && (sp->n <= 1 || 10 * sp->ascending >= 9 * (sp->n - 1));
    if ((sorted_list || (sp->kind == CF_FCI && gather && m->n >= 8192)) && make_sell(m.get(), true))
        m->auto64 = m->auto32 = SpmvChoice{7, 16, 2, {}, {}};
    *out = m.release();
    API_END
}

int cf_crs_destroy(cf_crs* m) {
    API_BEGIN
    if (m) {
        m->sp->ham->ctx->use();
        delete m;
    }
    API_END
}

static uint64_t count_rows(cf_crs* m, uint64_t r0, uint64_t r1, bool upper, std::vector<uint64_t>* per_row) {
    cf_ctx* c = m->sp->ham->ctx;
    ensure_crs(m);
    const uint64_t nr = r1 - r0;
    std::vector<uint64_t> cnt(nr);
    if (upper) {
        DBuf<uint64_t> d(nr);
        launch_count_upper(m->rowptr.p, m->col.p, r0, r1, d.p, c->stream);
        CF_HIP(hipMemcpyAsync(cnt.data(), d.p, nr * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
        CF_HIP(hipStreamSynchronize(c->stream));
    } else {
        std::vector<int64_t> rp(nr + 1);
        copy_sync(rp.data(), m->rowptr.p + r0, (nr + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
        for (uint64_t i = 0; i < nr; ++i) cnt[i] = (uint64_t)(rp[i + 1] - rp[i]);
    }
    uint64_t tot = 0;
    for (auto v : cnt) tot += v;
    if (per_row) per_row->swap(cnt);
    return tot;
}

static int export_rows(cf_crs* m, uint64_t r0, uint64_t r1, bool upper, uint64_t* row, uint32_t* col, double* val) {
    API_BEGIN
    CF_REQUIRE(m && row && col && val, "NULL argument");
    CF_REQUIRE(r0 <= r1 && r1 <= m->n, "row range out of bounds");
    cf_ctx* c = m->sp->ham->ctx;
    c->use();
    std::vector<uint64_t> per;
    uint64_t tot = count_rows(m, r0, r1, upper, &per);
    const uint64_t nr = r1 - r0;
    row[0] = 0;
    for (uint64_t i = 0; i < nr; ++i) row[i + 1] = row[i] + per[i];
    if (!nr) return CF_OK;
    DBuf<uint64_t> off(nr);
    copy_sync(off.p, row, nr * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream);
    DBuf<uint32_t> oc(tot ? tot : 1);
    DBuf<double> ov(tot ? tot : 1);
    launch_export_rows(m->rowptr.p, m->col.p, m->val.p, r0, r1, upper, off.p, oc.p, ov.p, c->stream);
    CF_HIP(hipMemcpyAsync(col, oc.p, tot * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipMemcpyAsync(val, ov.p, tot * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipStreamSynchronize(c->stream));
    API_END
}

int cf_crs_info(const cf_crs* m, uint64_t* n, uint64_t* nnz_full, uint64_t* nnz_upper) {
    API_BEGIN
    CF_REQUIRE(m, "matrix is NULL");
    cf_crs* mm = const_cast<cf_crs*>(m);
    if (nnz_full || nnz_upper) {  // counts of a stale (or deferred) CRS: build it first
        mm->sp->ham->ctx->use();
        ensure_crs(mm);
    }
    if (n) *n = m->n;
    if (nnz_full) *nnz_full = m->nnz;
    if (nnz_upper) {
        if (mm->nnz_upper < 0) {
            mm->sp->ham->ctx->use();
            mm->nnz_upper = (int64_t)count_rows(mm, 0, mm->n, true, nullptr);
        }
        *nnz_upper = (uint64_t)mm->nnz_upper;
    }
    API_END
}

int cf_crs_count_upper(cf_crs* m, uint64_t row_begin, uint64_t row_end, uint64_t* count) {
    API_BEGIN
    CF_REQUIRE(m && count, "NULL argument");
    CF_REQUIRE(row_begin <= row_end && row_end <= m->n, "row range out of bounds");
    m->sp->ham->ctx->use();
    *count = count_rows(m, row_begin, row_end, true, nullptr);
    API_END
}

int cf_crs_count_full(cf_crs* m, uint64_t row_begin, uint64_t row_end, uint64_t* count) {
    API_BEGIN
    CF_REQUIRE(m && count, "NULL argument");
    CF_REQUIRE(row_begin <= row_end && row_end <= m->n, "row range out of bounds");
    m->sp->ham->ctx->use();
    *count = count_rows(m, row_begin, row_end, false, nullptr);
    API_END
}

int cf_crs_export_upper(cf_crs* m, uint64_t row_begin, uint64_t row_end, uint64_t* row, uint32_t* col,
                        double* val) {
    return export_rows(m, row_begin, row_end, true, row, col, val);
}

int cf_crs_export_full(cf_crs* m, uint64_t row_begin, uint64_t row_end, uint64_t* row, uint32_t* col,
                       double* val) {
    return export_rows(m, row_begin, row_end, false, row, col, val);
}

int cf_crs_export_upper_u32(cf_crs* m, uint64_t row_begin, uint64_t row_end, uint32_t* row, uint32_t* col,
                            double* val) {
    uint64_t count = 0;
    int st = cf_crs_count_upper(m, row_begin, row_end, &count);
    if (st != CF_OK) return st;
    if (!row) return fail(CF_ERR_ARG, "NULL argument");
    if (count > UINT32_MAX)
        return fail(CF_ERR_UNSUPPORTED, "upper-triangle nonzeros " + std::to_string(count) +
                                            " exceed the 32-bit row pointer of SparseMatrix_CRS");
    try {
        std::vector<uint64_t> r64(row_end - row_begin + 1);
        st = export_rows(m, row_begin, row_end, true, r64.data(), col, val);
        if (st != CF_OK) return st;
        for (size_t i = 0; i < r64.size(); ++i) row[i] = (uint32_t)r64[i];
    } catch (const std::bad_alloc&) {
        return fail(CF_ERR_NOMEM, "host allocation failed");
    }
    return CF_OK;
}

int cf_mvprod(cf_crs* m, const double* x, double* y) {
    API_BEGIN
    CF_REQUIRE(m && x && y, "NULL argument");
    cf_ctx* c = m->sp->ham->ctx;
    c->use();
    ensure_staging(c, m->n);
    CF_HIP(hipMemcpyAsync(c->hx.p, x, m->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    spmv(m, CF_F64, c->hx.p, c->hy.p);
    CF_HIP(hipMemcpyAsync(y, c->hy.p, m->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipStreamSynchronize(c->stream));
    API_END
}

int cf_crs_spmv_bytes(const cf_crs* m, int value_prec, uint64_t* bytes, int* kernel) {
    API_BEGIN
    CF_REQUIRE(m && bytes && kernel, "NULL argument");
    const uint64_t sval = value_prec == CF_F32 ? 4 : 8;
    const SpmvChoice ch = resolve_spmv(m, value_prec);
    if (ch.kernel < 4 || ch.kernel == 7) {
        cf_crs* mm = const_cast<cf_crs*>(m);  // bytes of the CRS for the current Hamiltonian
        mm->sp->ham->ctx->use();
        ensure_crs(mm);
    }
    if (ch.kernel == 4) {
        const cf_space* sp = m->sp;
        const int Sb = sp->tb->S, Eb = sp->tb->E;
        *bytes = m->n * m->eargs.R * sval + 16 * m->n + (uint64_t)(Sb + Eb) * sp->db * 2;
    } else if (ch.kernel == 5) {
        // hop columns + string + diagonal + x once + y
        *bytes = m->n * ((uint64_t)m->dargs.np * m->dargs.nv * 4 + 32);
    } else if (ch.kernel == 6) {
        // diagonal + x once + y (the neighbour x blocks come back through L2)
        *bytes = m->n * 24;
    } else if (ch.kernel == 7) {
        // the stored entries (padding not counted) + chunk offsets + x once + y
        *bytes = m->nnz * (sval + 4) + ((m->n + 63) / 64 + 1) * 8 + 16 * m->n;
    } else {
        *bytes = m->nnz * (sval + 4) + (m->n + 1) * 8 + 16 * m->n;
    }
    *kernel = ch.kernel;
    API_END
}


int cf_crs_set_spmv_shape(cf_crs* m, int value_prec, int kernel, int p1, int p2) {
    API_BEGIN
    CF_REQUIRE(m, "matrix is NULL");
    CF_REQUIRE(value_prec == CF_F64 || value_prec == CF_F32, "value_prec must be CF_F64 or CF_F32");
    const SpmvChoice ch = make_choice(kernel, p1, p2, m->sp->db, value_prec);
    m->sp->ham->ctx->use();
    sync_ham(m);
    if (kernel == 5 && m->sp->kind == CF_DOCI && !m->doci_ready) make_doci_ell(m);
    if (kernel == 6 && m->sp->kind == CF_DOCI) {
        const int w = p1 > 0 ? p1 : doci_block_w(m->sp->ham->L);
        if (!m->blk_ready || m->bargs.w != w) make_doci_blocks(m, w);
    }
    check_shape(m, ch);
    if (kernel == 7) make_sell(m);
    (value_prec == CF_F32 ? m->auto32 : m->auto64) = ch;
    m->kernel = 0;
    API_END
}

int cf_crs_autotune(cf_crs* m, int value_prec, int reps, cf_tune_entry* table, int cap, int* count) {
    API_BEGIN
    CF_REQUIRE(m, "matrix is NULL");
    CF_REQUIRE(value_prec == CF_F64 || value_prec == CF_F32, "value_prec must be CF_F64 or CF_F32");
    CF_REQUIRE(value_prec == CF_F64 || m->want32, "matrix was built without FP32 values");
    CF_REQUIRE(reps >= 1 && reps <= 1000, "reps must be in [1, 1000]");
    cf_ctx* c = m->sp->ham->ctx;
    c->use();
    const uint64_t db = m->sp->db;
    ensure_crs(m);
    if (m->sp->kind == CF_DOCI && !m->doci_ready) make_doci_ell(m);
    if (m->sp->kind == CF_DOCI && !m->blk_ready && m->sp->ham->L - doci_block_w(m->sp->ham->L) <= 20)
        make_doci_blocks(m, doci_block_w(m->sp->ham->L));
    std::vector<SpmvChoice> cand;
    for (int l : {8, 16, 32, 64}) cand.push_back(make_choice(1, l, 0, db, value_prec));
    for (int l : {4, 8, 16, 32})
        for (int u : {1, 2}) cand.push_back(make_choice(3, l, u, db, value_prec));
    if (m->ell_ready) {
        std::vector<int> sizes = {256, 512, 768, 1024};
        for (int v : {1, 2, 4}) {
            if (db % v) continue;
            std::vector<int> ts = sizes;
            const uint64_t fit = (db + 64 * v - 1) / (64 * v) * 64;  // one chunk, fewest idle lanes
            if (fit <= 1024) ts.push_back((int)fit);
            std::sort(ts.begin(), ts.end());
            ts.erase(std::unique(ts.begin(), ts.end()), ts.end());
            for (int t : ts)
                if ((db + t - 1) / t <= 16) cand.push_back(make_choice(4, t, v, db, value_prec));
        }
        for (int t : {512, 1024}) cand.push_back(make_choice(4, t, 0, db, value_prec));
    }
    if (m->doci_ready)
        for (int t : {128, 256, 512})
            for (int u : {1, 2, 4, 8}) cand.push_back(make_choice(5, t, u, db, value_prec));
    if (m->blk_ready) cand.push_back(make_choice(6, 0, 0, db, value_prec));
    // the sliced copy doubles the matrix memory: timed when it exists (refilled
    // if new integrals made it stale), or for determinant lists (whose only
    // other kernels are the row gathers) when it fits in half of the free
    // device memory
    if ((m->scol.p || m->sp->kind == CF_DETS) && make_sell(m, true)) {
        for (int w : {1, 2, 4, 8, 16})
            for (int u : {1, 2, 4})
                if (u < 4 || w >= 4) cand.push_back(make_choice(7, w, u, db, value_prec));
    }
    ensure_staging(c, m->n);
    CF_HIP(hipMemsetAsync(c->hx.p, 0, m->n * sizeof(double), c->stream));
    hipEvent_t e0, e1;
    CF_HIP(hipEventCreate(&e0));
    CF_HIP(hipEventCreate(&e1));
    std::vector<double> ms(cand.size(), 0.0);
    try {
        for (size_t i = 0; i < cand.size(); ++i) {
            spmv_with(m, value_prec, cand[i], c->hx.p, c->hy.p);
            CF_HIP(hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps; ++r) spmv_with(m, value_prec, cand[i], c->hx.p, c->hy.p);
            CF_HIP(hipEventRecord(e1, c->stream));
            CF_HIP(hipEventSynchronize(e1));
            float t = 0.f;
            CF_HIP(hipEventElapsedTime(&t, e0, e1));
            ms[i] = (double)t / reps;
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    CF_HIP(hipEventDestroy(e0));
    CF_HIP(hipEventDestroy(e1));
    size_t best = 0;
    for (size_t i = 1; i < cand.size(); ++i)
        if (ms[i] < ms[best]) best = i;
    (value_prec == CF_F32 ? m->auto32 : m->auto64) = cand[best];
    m->kernel = 0;
    if (table)
        for (size_t i = 0; i < cand.size() && (int)i < cap; ++i) {
            const SpmvChoice& ch = cand[i];
            table[i].kernel = ch.kernel;
            table[i].p1 = ch.kernel == 4   ? ch.ell.threads
                          : ch.kernel == 5 ? ch.doci.threads
                          : ch.kernel == 6 ? m->bargs.w
                                           : ch.lanes;
            table[i].p2 = ch.kernel == 4 ? ch.ell.vec : ch.kernel == 5 ? ch.doci.sub : ch.unroll;
            table[i].ms = ms[i];
        }
    if (count) *count = (int)cand.size();
    API_END
}

int cf_mvprod_prec(cf_crs* m, int value_prec, const double* x, double* y) {
    API_BEGIN
    CF_REQUIRE(m && x && y, "NULL argument");
    CF_REQUIRE(value_prec == CF_F64 || value_prec == CF_F32, "value_prec must be CF_F64 or CF_F32");
    cf_ctx* c = m->sp->ham->ctx;
    c->use();
    ensure_staging(c, m->n);
    CF_HIP(hipMemcpyAsync(c->hx.p, x, m->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    spmv(m, value_prec, c->hx.p, c->hy.p);
    CF_HIP(hipMemcpyAsync(y, c->hy.p, m->n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipStreamSynchronize(c->stream));
    API_END
}

int cf_crs_set_spmv_kernel(cf_crs* m, int kind, int tile_rows) {
    API_BEGIN
    CF_REQUIRE(m, "matrix is NULL");
    CF_REQUIRE(kind >= 0 && kind <= 7, "kind must be 0, 1, 2, 3, 4, 5, 6 or 7");
    m->sp->ham->ctx->use();
    sync_ham(m);
    if (kind == 5) {
        CF_REQUIRE(m->sp->kind == CF_DOCI, "kernel 5 needs a DOCI matrix");
        if (tile_rows > 0)
            CF_REQUIRE(tile_rows % 64 == 0 && tile_rows <= 1024, "DOCI workgroup size must be a multiple of 64 <= 1024");
        if (!m->doci_ready) make_doci_ell(m);
        for (SpmvChoice* a : {&m->auto64, &m->auto32}) {
            if (a->kernel != 5) a->doci = doci_default_shape(m->n);
            if (tile_rows > 0) a->doci.threads = tile_rows;
        }
    }
    if (kind == 6) {
        CF_REQUIRE(m->sp->kind == CF_DOCI, "kernel 6 needs a DOCI matrix");
        const int w = tile_rows > 0 ? tile_rows : doci_block_w(m->sp->ham->L);
        if (!m->blk_ready || m->bargs.w != w) make_doci_blocks(m, w);
    }
    if (kind == 4) {
        if (tile_rows > 0)
            CF_REQUIRE(tile_rows % 64 == 0 && tile_rows <= 1024 && (m->sp->db + tile_rows - 1) / tile_rows <= 16,
                       "ELL workgroup size must be a multiple of 64 <= 1024 covering d_beta in 16 strides");
        if (!m->ell_ready) make_ell(m);
        for (SpmvChoice* a : {&m->auto64, &m->auto32}) {
            if (a->kernel != 4) a->ell = ell_default_shape(m->sp->db, a == &m->auto32 ? 4 : 8);
            if (tile_rows > 0) {
                a->ell.threads = tile_rows;
                if (a->ell.vec == 0) a->ell.vec = 1;
            }
        }
        CF_HIP(hipStreamSynchronize(m->sp->ham->ctx->stream));
    }
    if (kind == 7) {
        if (tile_rows > 0)
            CF_REQUIRE(tile_rows == 1 || tile_rows == 2 || tile_rows == 4 || tile_rows == 8 || tile_rows == 16,
                       "kernel 7: waves per chunk must be 1, 2, 4, 8 or 16");
        make_sell(m);
        m->sell_waves = tile_rows > 0 ? tile_rows : 16;
    }
    if (kind == 2) {
        CF_REQUIRE(m->sp->kind == CF_FCI, "the staged SpMV needs an FCI matrix");
        ensure_crs(m);
        const int W = tile_rows > 0 ? tile_rows : 16;
        make_fci_plan(m, W);
        CF_HIP(hipStreamSynchronize(m->sp->ham->ctx->stream));
    }
    m->kernel = kind;
    API_END
}

int cf_mvprod_dev(cf_crs* m, int value_prec, const double* dx, double* dy) {
    API_BEGIN
    CF_REQUIRE(m && dx && dy, "NULL argument");
    m->sp->ham->ctx->use();
    spmv(m, value_prec, dx, dy);
    API_END
}

void cf_solve_opts_default(cf_solve_opts* o) {
    if (!o) return;
    o->ncv = 42;
    o->maxrestart = 0;
    o->tol = 1e-12;
    o->precision = CF_F64;
    o->tol_f32 = 1e-8;
    o->seed = 20240601ULL;
    o->check_every = 0;
    o->tol_energy = 0.0;
    o->progress = nullptr;
    o->progress_user = nullptr;
    o->method = CF_LANCZOS;
}

}  // extern "C"

namespace {
// Shared driver of cf_solve / cf_solve_direct: FP64 thick-restart Lanczos, or
// an FP32-operator phase followed by the FP64 phase started from its Ritz
// vector.  The eigenvector stays resident in sp->vec.
// Preconditioned Davidson for the lowest eigenpair (E. R. Davidson, J. Comput.
// Phys. 17 (1975) 87): basis V and its image S = H V in HBM, the projected
// matrix G = V^T S on the host.  Each iteration: Ritz pair (theta, y) of G,
// x = V y and H x = S y in one pass, r = H x - theta x, correction
// t_i = r_i / (theta - H_ii), two Gram-Schmidt passes, one product.  A full
// basis restarts to its `keep` lowest Ritz vectors (V and S rotated by the
// same MFMA kernel as the Lanczos restart).
constexpr double kDgksEta = 0.717;

struct Davidson {
    cf_ctx* c;
    uint64_t n;
    int M, keep;
    const double* diag;
    DBuf<double> V, S, x, hx, t, partial, dg, Yd, dc, pnorm, pnorm2, rparts, tparts;
    DBuf<double> x1, r1parts, r1t, r1;  // second Ritz pair's residual (energy stop only)
    int nb;
    bool fused;  // small n: step kernels reduce their partials themselves
    int Gd;      // workgroups of the wide dot products
    double* hbuf;
    std::vector<double> G;  // M x M, row-major
    int m = 0;              // basis size
    int napply = 0, nrestart = 0;
    std::function<void(const double*, double*)> op;

    Davidson(cf_ctx* ctx, uint64_t n_, int maxb, const double* d) : c(ctx), n(n_), diag(d) {
        M = (int)std::min<uint64_t>((uint64_t)std::min(maxb, kDavMaxBasis), n);
        keep = std::max(1, std::min(4, M - 1));
        V.alloc((size_t)M * n);
        S.alloc((size_t)M * n);
        x.alloc(n);
        hx.alloc(n);
        t.alloc(n);
        nb = multi_dot_blocks(n);
        // the fused step's dot products run on at most 64 workgroups: measured
        // per class with 8 workers, fused against unfused: DOCI L=20 (184,756)
        // 339 / 306 solves/s, FCI L=11 (213,444) 81 / 89
        fused = n <= 200000;
        Gd = fused_dot_blocks(n);
        partial.alloc((size_t)64 * std::max(nb, Gd));
        dc.alloc(64);
        pnorm.alloc(256);
        pnorm2.alloc(256);
        rparts.alloc(256);
        tparts.alloc(256);
        dg.alloc(64);
        Yd.alloc(64 * 64);
        hbuf = c->pinned;
        G.assign((size_t)M * M, 0.0);
    }
    double* v(int i) { return V.p + (size_t)i * n; }
    double* s(int i) { return S.p + (size_t)i * n; }
    void dots(int k, const double* base, const double* w, double* out) {
        launch_multi_dot(n, k, base, w, partial.p, nb, c->stream);
        launch_reduce_partials(k, nb, partial.p, out, c->stream);
    }
    void add_column() {
        op(v(m), s(m));
        ++napply;
        dots(m + 1, V.p, s(m), dg.p);
        CF_HIP(hipMemcpyAsync(hbuf, dg.p, (m + 1) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CF_HIP(hipStreamSynchronize(c->stream));
        for (int i = 0; i <= m; ++i) G[(size_t)i * M + m] = G[(size_t)m * M + i] = hbuf[i];
        ++m;
    }
    // basis = the normalised vector in t
    void start_from_t() {
        m = 0;
        dots(1, t.p, t.p, dg.p + 63);
        CF_HIP(hipMemcpyAsync(hbuf, dg.p + 63, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CF_HIP(hipStreamSynchronize(c->stream));
        CF_REQUIRE(hbuf[0] > 0.0, "Davidson start vector is zero");
        launch_scale_copy(n, t.p, 1.0 / std::sqrt(hbuf[0]), v(0), c->stream);
        add_column();
    }
    // ||S y1 - theta1 V y1|| for the Ritz vector y1 of the first k basis
    // vectors: the residual of the second Ritz pair, which places a valid
    // lower end theta1 - ||r1|| of the Kato-Temple gap.  Uses hx and t (free
    // once the next basis vector is formed) and one more vector.
    double ritz_residual(int k, const double* y1, double theta1) {
        if (!x1.p) {
            x1.alloc(n);
            r1parts.alloc(256);
            r1t.alloc(256);
            r1.alloc(1);
        }
        launch_combine_residual(n, k, V.p, S.p, y1, theta1, 0.0, diag, x1.p, hx.p, t.p, r1parts.p, r1t.p, c->stream);
        launch_reduce_partials(1, fused_axpy_blocks(n), r1parts.p, r1.p, c->stream);
        double* hr = hbuf + 200;
        CF_HIP(hipMemcpyAsync(hr, r1.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        CF_HIP(hipStreamSynchronize(c->stream));
        return std::sqrt(std::max(0.0, hr[0]));
    }
    // Iterates to convergence; x holds the Ritz vector (|x| = 1) on return.
    // Returns 0 converged, 1 iteration limit, 2 breakdown, 3 stalled (with
    // stall > 0: no residual 10% below the best one for `stall` iterations,
    // the noise floor of a reduced-precision operator; the best one is reset
    // when the Ritz value falls by more than it).  One host wait per
    // iteration: the next basis vector and its product are formed before the
    // residual of the current Ritz vector is read back, so a converged solve
    // ends with one product that is not used.  The first `krylov` expansions
    // take the plain residual (a Krylov step) instead of the preconditioned
    // correction.
    int run(int maxit, int krylov, double tol, double tol_energy, int stall,
            const std::function<void(int, double, double)>& progress, double* theta_out, double* res_out) {
        std::vector<double> A, th, Y;
        double best = std::numeric_limits<double>::infinity(), theta_best = 0.0;
        int since_best = 0;
        double* ys = hbuf + 256;  // Ritz coefficients (<= 24), passed to the kernels by value
        double* Ys = hbuf + 320;  // staged restart rotation (<= 24 x 24)
        std::vector<double> y1;   // second Ritz vector's coefficients
        for (int it = 0;; ++it) {
            A.assign((size_t)m * m, 0.0);
            for (int i = 0; i < m; ++i)
                for (int j = 0; j < m; ++j) A[(size_t)i * m + j] = G[(size_t)i * M + j];
            sym_eig(m, A, th, Y);
            if (m == M) {  // restart on the keep lowest Ritz vectors
                std::memcpy(Ys, Y.data(), sizeof(double) * m * m);
                CF_HIP(hipMemcpyAsync(Yd.p, Ys, sizeof(double) * m * m, hipMemcpyHostToDevice, c->stream));
                launch_rotate_basis(n, m, keep, V.p, Yd.p, c->stream);
                launch_rotate_basis(n, m, keep, S.p, Yd.p, c->stream);
                std::fill(G.begin(), G.end(), 0.0);
                for (int i = 0; i < keep; ++i) G[(size_t)i * M + i] = th[i];
                m = keep;
                ++nrestart;
                for (int i = 0; i < m; ++i) ys[i] = i == 0 ? 1.0 : 0.0;
                y1.assign(m, 0.0);
                if (m >= 2) y1[1] = 1.0;
            } else {
                for (int i = 0; i < m; ++i) ys[i] = Y[(size_t)i * m + 0];
                y1.assign(m, 0.0);
                if (m >= 2)
                    for (int i = 0; i < m; ++i) y1[i] = Y[(size_t)i * m + 1];
            }
            const int mk = m;  // basis size of this Ritz step
            const double shift = it < krylov ? 0.0 : 1e-4;
            if (fused) {  // 8 launches besides the product
                launch_combine_residual(n, m, V.p, S.p, ys, th[0], shift, diag, x.p, hx.p, t.p, rparts.p, tparts.p,
                                        c->stream);
                launch_multi_dot_wide(n, m, V.p, t.p, partial.p, Gd, c->stream);
                launch_multi_axpy_r(n, m, Gd, partial.p, V.p, t.p, dc.p, pnorm.p, c->stream);
                // second pass only when the first removed more than half of t
                // (|t'| <= 0.717 |t|, the refinement test ARPACK's dsaitr uses)
                launch_multi_dot_wide(n, m, V.p, t.p, partial.p, Gd, c->stream, tparts.p, pnorm.p, kDgksEta);
                launch_multi_axpy_r(n, m, Gd, partial.p, V.p, t.p, dc.p, pnorm2.p, c->stream, tparts.p, pnorm.p,
                                    kDgksEta);
                launch_scale_rnorm_parts(n, t.p, pnorm.p, rparts.p, v(m), dg.p + 62, c->stream, tparts.p, pnorm2.p,
                                         kDgksEta);
                op(v(m), s(m));
                ++napply;
                launch_multi_dot_wide(n, m + 1, V.p, s(m), partial.p, Gd, c->stream);
                launch_reduce_partials(m + 1, Gd, partial.p, dg.p, c->stream);
            } else {  // two passes over V (and one over S) before the product
                // x, t, first-pass dots V^T t -> dg[0, m), |t|^2 -> dg[61], |r|^2 -> dg[62]
                launch_combine_res_dot(n, m, V.p, S.p, ys, th[0], shift, diag, x.p, t.p, partial.p, nb, 61, 62,
                                       c->stream);
                launch_reduce_partials(m, nb, partial.p, dg.p, c->stream);
                launch_reduce_partials(2, nb, partial.p + (size_t)61 * nb, dg.p + 61, c->stream);
                // t -= V dg[0, m); second-pass dots -> dg[0, m), |t'|^2 -> dg[m]
                launch_axpy_dot(n, m, V.p, dg.p, t.p, partial.p, nb, m, c->stream);
                launch_reduce_partials(m + 1, nb, partial.p, dg.p, c->stream);
                // second pass only when the first removed more than half of t
                launch_multi_axpy(n, m, V.p, dg.p, t.p, c->stream, dg.p + 61, dg.p + m, kDgksEta);
                dots(1, t.p, t.p, dg.p + 63);
                launch_scale_rnorm(n, t.p, dg.p + 63, v(m), c->stream);
                op(v(m), s(m));
                ++napply;
                dots(m + 1, V.p, s(m), dg.p);
            }
            CF_HIP(hipMemcpyAsync(hbuf, dg.p, 64 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            CF_HIP(hipStreamSynchronize(c->stream));
            const double res = std::sqrt(std::max(0.0, hbuf[62]));
            *theta_out = th[0];
            *res_out = res;
            if (progress) progress(napply, th[0], res);
            if (res <= tol * std::fabs(th[0])) return 0;
            // Kato-Temple: the optimistic gap theta1 - theta0 first, then the
            // safeguarded one with the second pair's residual (one more pass)
            if (tol_energy > 0.0 && mk >= 2 && th.size() >= 2 && res * res <= tol_energy * (th[1] - th[0]) &&
                temple_accept(res, th[0], th[1], ritz_residual(mk, y1.data(), th[1]), tol_energy))
                return 0;
            if (!(hbuf[63] > 1e-28)) return 2;
            if (it >= maxit) return 1;
            // a Ritz value that fell by more than the best residual belongs
            // to another eigenvalue (|theta - lambda| <= |r|): measure the
            // stall from there
            if (res < 0.9 * best || th[0] < theta_best - best) {
                best = res;
                theta_best = th[0];
                since_best = 0;
            } else if (stall > 0 && ++since_best >= stall) {
                return 3;
            }
            for (int i = 0; i <= m; ++i) G[(size_t)i * M + m] = G[(size_t)m * M + i] = hbuf[i];
            ++m;
        }
    }
};

void run_davidson(cf_space* sp, uint64_t n, const cf_solve_opts& o, std::function<void(const double*, double*)> op64,
                  std::function<void(const double*, double*)> op32, double* eigvec, cf_solve_info* info) {
    cf_ctx* c = sp->ham->ctx;
    CF_REQUIRE(sp->diag.p != nullptr, "the Davidson solver needs a space with a diagonal (FCI or DOCI)");
    const auto t0 = std::chrono::steady_clock::now();
    Davidson d(c, n, std::min(o.ncv, 24), sp->diag.p);
    const int maxit = o.maxrestart > 0 ? o.maxrestart * d.M : 20000;
    // Start: the deterministic random vector of the Lanczos path, expanded by
    // plain Krylov steps before the preconditioned corrections begin.  A start
    // on the lowest-diagonal determinants converges fast but stays in their
    // symmetry sector (H conserves S^2; a closed-shell determinant is a pure
    // singlet), and it can settle on a singlet above a triplet ground state;
    // the Krylov steps from a random vector, as in Lanczos, let the lowest
    // state of any sector lead before the diagonal preconditioner sharpens it.
    constexpr int kKrylovSteps = 8;
    launch_fill_start(n, o.seed, d.t.p, c->stream);
    int nmv32 = 0, status = 0;
    double theta = 0.0, res = 0.0;
    auto prog = [&](int nmv, double th, double r) {
        if (o.progress) o.progress(o.progress_user, nmv, th, r);
    };
    if (o.precision == CF_F32) {
        CF_REQUIRE((bool)op32, "the FP32 operator is not available (build with CF_F32)");
        d.op = op32;
        d.start_from_t();
        // the FP32 residual bottoms out near the operator's rounding (the
        // GEMM parts accumulate in FP32): stop this phase when it stalls
        d.run(maxit, kKrylovSteps, o.tol_f32, o.tol_energy, 24, prog, &theta, &res);
        nmv32 = d.napply;
        launch_copy(n, d.x.p, d.t.p, c->stream);  // polish from the FP32 Ritz vector
        d.napply = 0;
    }
    d.op = op64;
    d.start_from_t();
    auto prog64 = [&](int nmv, double th, double r) {
        if (o.progress) o.progress(o.progress_user, nmv32 + nmv, th, r);
    };
    // FP64: a safety stop only (200 iterations without a 10% better residual)
    status = d.run(maxit, o.precision == CF_F32 ? 0 : kKrylovSteps, o.tol, o.tol_energy, 200, prog64, &theta, &res);
    info->eigval = theta;
    info->energy = theta + sp->ham->econst;
    info->resid = res;
    info->n_matvec = d.napply;
    info->n_matvec_f32 = nmv32;
    info->n_restart = d.nrestart;
    info->status = status == 0 ? 0 : 1;
    keep_eigvec(sp, n, d.x.p, eigvec);
    info->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void run_solve(cf_space* sp, uint64_t n, const cf_solve_opts* opts, std::function<void(const double*, double*)> op64,
               std::function<void(const double*, double*)> op32, double* eigvec, cf_solve_info* info) {
    cf_solve_opts o;
    if (opts) o = *opts;
    else cf_solve_opts_default(&o);
    CF_REQUIRE(o.ncv >= 2 && o.ncv <= 63, "ncv must be in [2, 63]");
    CF_REQUIRE(o.tol > 0.0, "tol must be > 0");
    CF_REQUIRE(o.check_every >= 0, "check_every must be >= 0");
    CF_REQUIRE(o.tol_energy >= 0.0, "tol_energy must be >= 0");
    CF_REQUIRE(o.precision == CF_F64 || o.precision == CF_F32, "precision must be CF_F64 or CF_F32");
    CF_REQUIRE(o.method == CF_LANCZOS || o.method == CF_DAVIDSON || o.method == CF_SOLVER_AUTO,
               "method must be CF_LANCZOS, CF_DAVIDSON or CF_SOLVER_AUTO");
    // measured on MI355X, FP64 solves to tol 1e-12: FCI L=12 6a6b 476 -> 417 ms,
    // DOCI L=24 12 pairs 136 -> 111 ms with Davidson; below a few thousand
    // determinants the solves are launch-bound and Lanczos is as fast
    if (o.method == CF_SOLVER_AUTO) o.method = (sp->diag.p && n >= 2000) ? CF_DAVIDSON : CF_LANCZOS;
    if (o.method == CF_DAVIDSON) {
        run_davidson(sp, n, o, op64, op32, eigvec, info);
        return;
    }
    cf_ctx* c = sp->ham->ctx;
    const auto t0 = std::chrono::steady_clock::now();
    const int maxr = o.maxrestart > 0 ? o.maxrestart : (int)std::min<uint64_t>(3 * n, 1u << 30);
    DeviceBackend b(c, n, o.ncv);
    int nmv32 = 0, nrs = 0;
    bool keep = false;
    if (o.precision == CF_F32) {
        CF_REQUIRE((bool)op32, "the FP32 operator is not available (build with CF_F32)");
        b.op = op32;
        ThickRestartLanczos sm(&b, n, o.ncv, o.tol_f32, maxr, o.seed, false, o.check_every, o.tol_energy);
        if (o.progress)
            sm.set_progress([&o](int nmv, double th, double r) { o.progress(o.progress_user, nmv, th, r); });
        while (sm.step() == ThickRestartLanczos::NEED_APPLY) b.apply(sm.pending());
        nmv32 = sm.n_apply();
        nrs += sm.n_restart();
        keep = true;  // v0 holds the FP32-phase Ritz vector: polish from it
    }
    b.op = op64;
    // the polish starts from a vector converged to tol_f32: test every few
    // steps instead of only when the basis is full
    const int check = keep ? (o.check_every > 0 ? o.check_every : 4) : o.check_every;
    ThickRestartLanczos sm(&b, n, o.ncv, o.tol, maxr, o.seed, keep, check, o.tol_energy);
    if (o.progress)
        sm.set_progress([&o, nmv32](int nmv, double th, double r) { o.progress(o.progress_user, nmv32 + nmv, th, r); });
    while (sm.step() == ThickRestartLanczos::NEED_APPLY) b.apply(sm.pending());
    info->eigval = sm.eigval();
    info->energy = sm.eigval() + sp->ham->econst;
    info->resid = sm.resid();
    info->n_matvec = sm.n_apply();
    info->n_matvec_f32 = nmv32;
    info->n_restart = nrs + sm.n_restart();
    info->status = sm.status();
    keep_eigvec(sp, n, b.v(0), eigvec);
    info->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
}  // namespace

namespace {
#define CF_BLAS(call)                                                                       \
    do {                                                                                    \
        rocblas_status st_ = (call);                                                        \
        if (st_ != rocblas_status_success)                                                  \
            throw ::cf::Error(::cf::ERR_HIP, std::string(#call) + ": rocblas status " +     \
                                                 std::to_string((int)st_));                 \
    } while (0)

// y = Ha C + C Hb^T + sigma_ab.  The same-spin parts are two FP64 GEMMs with
// the dense string Hamiltonians for either operand precision (measured: a
// kernel applying the excitation lists directly, with FP64 accumulation, was
// 4-6x slower than the library GEMM at L = 11 and 16); the alpha-beta kernel
// takes FP32 or FP64 operands and sums in FP64.  Row-major C (da x db) is the
// column-major db x da matrix Cc, so y_c = Cc Ha + Hb Cc (Ha, Hb symmetric).
// CI_Big::mvprod (CIMethod.cpp:943-984): the stored leading block's product
// into y[0, m), then every other element on the fly.
void dets_sigma_apply(cf_sigma* S, const double* x, double* y) {
    cf_space* sp = S->sp;
    cf_ctx* c = sp->ham->ctx;
    KTimer t(c, "sigma_dets");
    if (S->block) spmv(S->block, CF_F64, x, y);
    DetLookupArgs a{};
    a.r.n = sp->n;
    a.r.nup = sp->ham->nup;
    a.r.ndown = sp->ham->ndown;
    a.r.up = sp->dup.p;
    a.r.down = sp->ddown.p;
    a.r.H = sp->ham->dev();
    a.skey = S->skey.p;
    a.sidx = S->sidx.p;
    a.L = sp->ham->L;
    launch_det_sigma_lookup(a, x, y, S->stored, c->stream);
}

void sigma_apply(cf_sigma* S, int prec, const double* x, double* y) {
    cf_space* sp = S->sp;
    CF_REQUIRE(S->version == sp->version,
               "the space's Hamiltonian changed after this sigma was created: create the sigma again");
    if (S->dets) {
        dets_sigma_apply(S, x, y);  // FP64 elements evaluated on the fly in either precision
        return;
    }
    cf_ctx* c = sp->ham->ctx;
    const int m = (int)sp->db, nn = (int)sp->da;
    KTimer t(c, prec == CF_F32 ? "sigma_f32" : "sigma_f64");
    const double one = 1.0, zero = 0.0;
    CF_BLAS(rocblas_dgemm(S->blas, rocblas_operation_none, rocblas_operation_none, m, nn, nn, &one, x, m, S->Ha.p,
                          nn, &zero, y, m));
    CF_BLAS(rocblas_dgemm(S->blas, rocblas_operation_none, rocblas_operation_none, m, nn, m, &one, S->Hb, m, x, m,
                          &one, y, m));
    if (prec == CF_F32) {
        CF_REQUIRE(S->f32, "sigma plan was created without FP32 operands");
        launch_to_f32(sp->n, x, S->x32.p, c->stream);
        if (S->shape.kernel != 3 || !launch_sigma_ab3<float>(S->args, sp->ham->ndown, S->x32.p, y, c->stream))
            launch_sigma_ab2<float>(S->args, S->shape, S->x32.p, y, c->stream);
    } else {
        if (S->shape.kernel != 3 || !launch_sigma_ab3<double>(S->args, sp->ham->ndown, x, y, c->stream))
            launch_sigma_ab2<double>(S->args, S->shape, x, y, c->stream);
    }
}
}  // namespace

extern "C" {

int cf_solve(cf_crs* m, const cf_solve_opts* opts, double* eigvec, cf_solve_info* info) {
    API_BEGIN
    CF_REQUIRE(m && info, "NULL argument");
    m->sp->ham->ctx->use();
    std::function<void(const double*, double*)> op32;
    if (m->want32) op32 = [m](const double* x, double* y) { spmv(m, CF_F32, x, y); };
    run_solve(m->sp, m->n, opts, [m](const double* x, double* y) { spmv(m, CF_F64, x, y); }, op32, eigvec, info);
    API_END
}

int cf_space_eigvec(cf_space* sp, double* eigvec) {
    API_BEGIN
    CF_REQUIRE(sp && eigvec, "NULL argument");
    CF_REQUIRE(sp->vec_valid, "no solve has run on this space");
    sp->ham->ctx->use();
    copy_sync(eigvec, sp->vec.p, sp->n * sizeof(double), hipMemcpyDeviceToHost, sp->ham->ctx->stream);
    if (sp->vec_sign_pending) {
        std::vector<double> hv(eigvec, eigvec + sp->n);
        sign_fix(hv);
        std::memcpy(eigvec, hv.data(), sp->n * sizeof(double));
    }
    API_END
}

int cf_sigma_create(cf_space* sp, int prec, cf_sigma** out) {
    API_BEGIN
    CF_REQUIRE(sp && out, "NULL argument");
    CF_REQUIRE(sp->kind == CF_FCI, "the on-the-fly sigma is for FCI spaces");
    CF_REQUIRE(prec == CF_F64 || prec == CF_F32, "prec must be CF_F64 or CF_F32");
    CF_REQUIRE(sp->ham->L <= 31 && sp->db < (1ull << 21) && sp->tb->S - 1 <= 64,
               "on-the-fly sigma supports L <= 31, beta dimension < 2^21, <= 64 beta singles");
    CF_REQUIRE(sp->da <= 2147483647ull && sp->db <= 2147483647ull, "dimension too large for the GEMMs");
    cf_ctx* c = sp->ham->ctx;
    c->use();
    std::unique_ptr<cf_sigma> S(new cf_sigma());
    S->sp = sp;
    S->version = sp->version;
    S->f32 = prec == CF_F32;
    S->Ha.alloc(sp->da * sp->da);
    launch_dense_string_ham<double>(sp->da, sp->ta->E, sp->ta->full.p, S->Ha.p, c->stream);
    if (sp->tb != sp->ta.get()) {
        S->Hb_own.alloc(sp->db * sp->db);
        launch_dense_string_ham<double>(sp->db, sp->tb->E, sp->tb->full.p, S->Hb_own.p, c->stream);
        S->Hb = S->Hb_own.p;
    } else {
        S->Hb = S->Ha.p;
    }
    if (S->f32) S->x32.alloc(sp->n);
    S->blas = c->blas();
    // measured on MI355X: 256 beta strings per workgroup up to d_beta = 512 (L = 11:
    // 176 us against 187 us at 512), 512 up to L = 14 (d_beta <= 3432), 768 above
    // (L = 16: 771 ms against 1458 ms at 1024); d_beta <= 128 in one 128-thread chunk
    S->shape.threads = sp->db <= 128 ? 128 : sp->db <= 512 ? 256 : sp->db <= 4096 ? 512 : 768;
    SigmaArgs& a = S->args;
    a.da = sp->da;
    a.db = sp->db;
    a.L = sp->ham->L;
    a.Sa = sp->ta->S;
    a.Sb = sp->tb->S;
    CF_REQUIRE(sp->da * ((sp->db + 255) / 256) < (1ull << 31), "too many sigma workgroups");
    a.sa = sp->ta->s.p;
    a.sb = sp->tb->s.p;
    a.suba = sp->ta->sub.p;
    a.subb = sp->tb->sub.p;
    a.V = sp->ham->dV.p;
    CF_REQUIRE(sigma_ab_lds_bytes(a, false) <= 160 * 1024, "beta string space too large for the LDS row");
    const uint64_t L2 = (uint64_t)a.L * a.L;
    S->vpair64.alloc(L2 * L2);
    S->vself64.alloc(sp->da * L2);
    launch_sigma_tables<double>(a, S->vpair64.p, S->vself64.p, c->stream);
    a.vpair64 = S->vpair64.p;
    a.vself64 = S->vself64.p;
    if (S->f32) {
        S->vpair32.alloc(L2 * L2);
        S->vself32.alloc(sp->da * L2);
        launch_sigma_tables<float>(a, S->vpair32.p, S->vself32.p, c->stream);
        a.vpair32 = S->vpair32.p;
        a.vself32 = S->vself32.p;
    }
    CF_HIP(hipStreamSynchronize(c->stream));
    *out = S.release();
    API_END
}

int cf_sigma_set_threads(cf_sigma* s, int threads) {
    API_BEGIN
    CF_REQUIRE(s, "NULL argument");
    CF_REQUIRE(threads == 128 || threads == 256 || threads == 512 || threads == 768 || threads == 1024,
               "threads must be 128, 256, 512, 768 or 1024");
    s->shape.threads = threads;
    API_END
}

int cf_sigma_set_kernel(cf_sigma* s, int kernel) {
    API_BEGIN
    CF_REQUIRE(s, "NULL argument");
    CF_REQUIRE(kernel == 1 || kernel == 2 || kernel == 3, "kernel must be 1, 2 or 3");
    s->shape.kernel = kernel;
    API_END
}

int cf_sigma_destroy(cf_sigma* s) {
    API_BEGIN
    if (s) {
        s->sp->ham->ctx->use();
        delete s->block;
        delete s->block_sp;
        delete s;
    }
    API_END
}

int cf_dets_sigma_create(cf_space* sp, uint64_t stored_rows, cf_sigma** out) {
    API_BEGIN
    CF_REQUIRE(sp && out, "NULL argument");
    CF_REQUIRE(sp->kind == CF_DETS, "cf_dets_sigma_create is for determinant-list spaces");
    CF_REQUIRE(sp->ham->L <= 32, "the on-the-fly determinant-list product supports L <= 32");
    CF_REQUIRE(stored_rows <= sp->n, "stored_rows exceeds the number of determinants");
    cf_ctx* c = sp->ham->ctx;
    c->use();
    std::unique_ptr<cf_sigma> S(new cf_sigma());
    S->sp = sp;
    S->version = sp->version;
    S->dets = true;
    S->stored = stored_rows;
    S->skey.alloc(sp->n);
    S->sidx.alloc(sp->n);
    sort_det_keys(sp->n, sp->dup.p, sp->ddown.p, S->skey.p, S->sidx.p, 32 + sp->ham->L, c->stream);
    if (stored_rows > 0) {  // the leading block as its own list, built like FCI_File's matrix
        std::unique_ptr<cf_space> b(new cf_space());
        b->ham = sp->ham;
        b->kind = CF_DETS;
        b->n = b->da = stored_rows;
        b->db = 1;
        b->version = sp->version;
        b->dup.alloc(stored_rows);
        b->ddown.alloc(stored_rows);
        CF_HIP(hipMemcpyAsync(b->dup.p, sp->dup.p, stored_rows * sizeof(str_t), hipMemcpyDeviceToDevice, c->stream));
        CF_HIP(hipMemcpyAsync(b->ddown.p, sp->ddown.p, stored_rows * sizeof(str_t), hipMemcpyDeviceToDevice,
                              c->stream));
        std::unique_ptr<cf_crs> m(new cf_crs());
        m->sp = b.get();
        m->n = stored_rows;
        m->zero = 1e-14;  // the reference's _zero (SparseMatrix_CRS.h:116)
        m->version = b->version;
        build_crs_arrays(m.get());
        // the row-gather choice of cf_crs_build for FP64 values
        m->auto64.kernel = m->auto32.kernel = 1;
        m->auto64.lanes = m->auto32.lanes = m->lpr;
        if (m->n && (double)m->nnz / (double)m->n >= 512) m->auto64 = SpmvChoice{3, 16, 2, {}, {}};
        S->block_sp = b.release();
        S->block = m.release();
    }
    CF_HIP(hipStreamSynchronize(c->stream));
    *out = S.release();
    API_END
}

int cf_sigma_apply_dev(cf_sigma* s, int value_prec, const double* dx, double* dy) {
    API_BEGIN
    CF_REQUIRE(s && dx && dy, "NULL argument");
    s->sp->ham->ctx->use();
    sigma_apply(s, value_prec, dx, dy);
    API_END
}

int cf_sigma_apply(cf_sigma* s, int value_prec, const double* x, double* y) {
    API_BEGIN
    CF_REQUIRE(s && x && y, "NULL argument");
    cf_ctx* c = s->sp->ham->ctx;
    c->use();
    const uint64_t n = s->sp->n;
    ensure_staging(c, n);
    CF_HIP(hipMemcpyAsync(c->hx.p, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    sigma_apply(s, value_prec, c->hx.p, c->hy.p);
    CF_HIP(hipMemcpyAsync(y, c->hy.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipStreamSynchronize(c->stream));
    API_END
}

int cf_solve_direct(cf_sigma* s, const cf_solve_opts* opts, double* eigvec, cf_solve_info* info) {
    API_BEGIN
    CF_REQUIRE(s && info, "NULL argument");
    s->sp->ham->ctx->use();
    std::function<void(const double*, double*)> op32;
    if (s->f32) op32 = [s](const double* x, double* y) { sigma_apply(s, CF_F32, x, y); };
    run_solve(s->sp, s->sp->n, opts, [s](const double* x, double* y) { sigma_apply(s, CF_F64, x, y); }, op32, eigvec,
              info);
    API_END
}

int cf_rdm1_direct(cf_space* sp, const double* cvec, double* one) {
    API_BEGIN
    CF_REQUIRE(sp && cvec && one, "NULL argument");
    CF_REQUIRE(sp->kind == CF_FCI, "cf_rdm1_direct is for FCI spaces");
    cf_ctx* c = sp->ham->ctx;
    c->use();
    const int L = sp->ham->L;
    const uint64_t n = sp->n, da = sp->da, db = sp->db;
    DBuf<double> C(n), Ct(n), da_out(da * (uint64_t)sp->ta->S), db_out(db * (uint64_t)sp->tb->S);
    CF_HIP(hipMemcpyAsync(C.p, cvec, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    {
        KTimer t(c, "rdm");
        launch_row_dots(da, db, sp->ta->S, sp->ta->sub.p, C.p, da_out.p, c->stream);
        launch_transpose(da, db, C.p, Ct.p, c->stream);
        launch_row_dots(db, da, sp->tb->S, sp->tb->sub.p, Ct.p, db_out.p, c->stream);
    }
    std::vector<double> ha(da * (uint64_t)sp->ta->S), hb(db * (uint64_t)sp->tb->S);
    std::vector<Exc> la(da * (uint64_t)sp->ta->S), lb(db * (uint64_t)sp->tb->S);
    CF_HIP(hipMemcpyAsync(ha.data(), da_out.p, ha.size() * 8, hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipMemcpyAsync(hb.data(), db_out.p, hb.size() * 8, hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipMemcpyAsync(la.data(), sp->ta->sub.p, la.size() * sizeof(Exc), hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipMemcpyAsync(lb.data(), sp->tb->sub.p, lb.size() * sizeof(Exc), hipMemcpyDeviceToHost, c->stream));
    std::vector<str_t> sa(da), sb(db);
    CF_HIP(hipMemcpyAsync(sa.data(), sp->ta->s.p, da * 8, hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipMemcpyAsync(sb.data(), sp->tb->s.p, db * 8, hipMemcpyDeviceToHost, c->stream));
    CF_HIP(hipStreamSynchronize(c->stream));
    // gamma^s_pr = sum over strings K, entries e of K's list: self -> n_p(K) |C_K|^2 on the
    // diagonal; single p -> r: sign * C_K . C_J at (p, r) (= <a+_p a_r>, symmetric)
    std::vector<double> g(2 * (size_t)L * L, 0.0);
    for (int s = 0; s < 2; ++s) {
        const uint64_t d = s == 0 ? da : db;
        const int S = s == 0 ? sp->ta->S : sp->tb->S;
        const std::vector<double>& dots = s == 0 ? ha : hb;
        const std::vector<Exc>& lst = s == 0 ? la : lb;
        const std::vector<str_t>& strs = s == 0 ? sa : sb;
        double* gs = g.data() + (size_t)s * L * L;
        for (uint64_t K = 0; K < d; ++K)
            for (int e = 0; e < S; ++e) {
                const Exc& x = lst[K * S + e];
                const double v = dots[K * S + e];
                if (x.type == 0) {
                    for (str_t t = strs[K]; t; t &= t - 1) gs[__builtin_ctzll(t) * (L + 1)] += v;
                } else {
                    gs[x.f * L + x.t] += x.sign < 0 ? -v : v;
                }
            }
    }
    const size_t osz = (size_t)L * (L + 1) / 2;
    for (int s = 0; s < 2; ++s)
        for (int p = 0; p < L; ++p)
            for (int q = p; q < L; ++q)
                one[s * osz + (size_t)p * L + q - (size_t)p * (p + 1) / 2] = g[(size_t)s * L * L + p * L + q];
    API_END
}

int cf_rotate_integrals(cf_ctx* ctx, int L, int batch, const double* h, const double* V, const double* U,
                        double* h_out, double* V_out) {
    API_BEGIN
    CF_REQUIRE(ctx && h && V && U && h_out && V_out, "NULL argument");
    CF_REQUIRE(L >= 1 && L <= kMaxL && batch >= 1, "bad L / batch");
    ctx->use();
    const size_t L2 = (size_t)L * L, L4 = L2 * L2;
    DBuf<double> dh(batch * L2), dV(batch * L4), dU(batch * L2), oh(batch * L2), oV(batch * L4), tmp(batch * L4);
    CF_HIP(hipMemcpyAsync(dh.p, h, dh.bytes(), hipMemcpyHostToDevice, ctx->stream));
    CF_HIP(hipMemcpyAsync(dV.p, V, dV.bytes(), hipMemcpyHostToDevice, ctx->stream));
    CF_HIP(hipMemcpyAsync(dU.p, U, dU.bytes(), hipMemcpyHostToDevice, ctx->stream));
    {
        KTimer t(ctx, "rotate");
        rotate_integrals_dev(L, batch, dh.p, dV.p, dU.p, oh.p, oV.p, tmp.p, ctx->stream);
    }
    CF_HIP(hipMemcpyAsync(h_out, oh.p, oh.bytes(), hipMemcpyDeviceToHost, ctx->stream));
    CF_HIP(hipMemcpyAsync(V_out, oV.p, oV.bytes(), hipMemcpyDeviceToHost, ctx->stream));
    CF_HIP(hipStreamSynchronize(ctx->stream));
    API_END
}

int cf_ham_rotate(const cf_ham* src, const double* U, cf_ham** out) {
    API_BEGIN
    CF_REQUIRE(src && U && out, "NULL argument");
    cf_ctx* ctx = src->ctx;
    ctx->use();
    const int L = src->L;
    const size_t L2 = (size_t)L * L, L3 = L2 * L, L4 = L3 * L;
    std::unique_ptr<cf_ham> H(new cf_ham());
    H->ctx = ctx;
    H->L = L;
    H->nup = src->nup;
    H->ndown = src->ndown;
    H->econst = src->econst;
    H->dh.alloc(L2);
    H->dV.alloc(L4);
    H->dW.alloc(L3);
    H->dJm.alloc(L2);
    DBuf<double> dU(L2), tmp(L4);
    CF_HIP(hipMemcpyAsync(dU.p, U, L2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    {
        KTimer t(ctx, "rotate");
        rotate_integrals_dev(L, 1, src->dh.p, src->dV.p, dU.p, H->dh.p, H->dV.p, tmp.p, ctx->stream);
    }
    launch_derived(L, H->dV.p, H->dW.p, H->dJm.p, ctx->stream);
    CF_HIP(hipStreamSynchronize(ctx->stream));  // U is the caller's host array
    *out = H.release();
    API_END
}

int cf_ham_get(const cf_ham* ham, double* h, double* V) {
    API_BEGIN
    CF_REQUIRE(ham, "ham is NULL");
    ham->ctx->use();
    const size_t L2 = (size_t)ham->L * ham->L;
    if (h) copy_sync(h, ham->dh.p, L2 * sizeof(double), hipMemcpyDeviceToHost, ham->ctx->stream);
    if (V) copy_sync(V, ham->dV.p, L2 * L2 * sizeof(double), hipMemcpyDeviceToHost, ham->ctx->stream);
    API_END
}

int cf_rdm(cf_space* sp, const double* cvec, double* one, double* two) {
    API_BEGIN
    CF_REQUIRE(sp && cvec && one, "NULL argument");
    cf_ctx* c = sp->ham->ctx;
    c->use();
    const int L = sp->ham->L;
    std::vector<double> g(2 * L * L), G;
    if (two) G.resize((size_t)3 * L * L * L * L);
    if (sp->kind == CF_DETS) {
        KTimer t(c, "rdm");
        rdm_dets(c->stream, L, sp->n, sp->dup.p, sp->ddown.p, cvec, g.data(), two ? G.data() : nullptr);
    } else {
        KTimer t(c, "rdm");
        rdm_dense(c->stream, sp->kind, L, sp->ham->nup, sp->ham->ndown, sp->da, sp->db, sp->ta->s.p, sp->tb->s.p,
                  sp->ta->sub.p, sp->tb->sub.p, sp->ta->S, sp->tb->S, c->binom.p, cvec, sp->n, g.data(),
                  two ? G.data() : nullptr);
    }
    pack_rdm(L, g.data(), two ? G.data() : nullptr, one, two);
    API_END
}

// Batched small DOCI solves (one workgroup per request, see batch.hip).
int cf_solve_doci_batch(cf_ctx* ctx, int count, const cf_request* reqs, cf_result* results) {
    API_BEGIN
    CF_REQUIRE(ctx && (count == 0 || (reqs && results)), "NULL argument");
    CF_REQUIRE(count >= 0, "negative count");
    if (!count) return CF_OK;
    ctx->use();
    hipStream_t st = ctx->stream;
    const int cap = doci_batch_capacity();
    cf_solve_opts dflt;
    cf_solve_opts_default(&dflt);
    std::vector<int> ok(count, 0);
    for (int i = 0; i < count; ++i) {
        const cf_request& r = reqs[i];
        cf_result& o = results[i];
        std::memset(&o, 0, sizeof(o));
        o.tag = r.tag;
        o.status = CF_ERR_ARG;
        const char* why = nullptr;
        if (r.kind != CF_DOCI) why = "batched solves are DOCI only";
        else if (!r.h || !r.V) why = "NULL integrals";
        else if (r.L < 1 || r.L > kBatchMaxL || r.nup != r.ndown || r.nup < 1 || r.nup >= r.L)
            why = "batched DOCI needs 1 <= nalpha == nbeta < L <= 32";
        else if (binom_host(r.L, r.nup) > (uint64_t)cap) why = "DOCI space too large for a batched solve";
        if (why) {
            std::snprintf(o.error, sizeof(o.error), "%s", why);
            continue;
        }
        ok[i] = 1;
        o.n = binom_host(r.L, r.nup);
    }
    // integrals: per L, one batched rotation (requests without U take the identity)
    std::map<int, std::vector<int>> byL;
    for (int i = 0; i < count; ++i)
        if (ok[i]) byL[reqs[i].L].push_back(i);
    std::vector<DBuf<double>> hbufs, vbufs;
    std::vector<const double*> hptr(count, nullptr), vptr(count, nullptr);
    for (auto& g : byL) {
        const int L = g.first, B = (int)g.second.size();
        const size_t L2 = (size_t)L * L, L4 = L2 * L2;
        std::vector<double> hh(B * L2), vv(B * L4), uu(B * L2, 0.0);
        for (int b = 0; b < B; ++b) {
            const cf_request& r = reqs[g.second[b]];
            std::memcpy(&hh[b * L2], r.h, L2 * sizeof(double));
            std::memcpy(&vv[b * L4], r.V, L4 * sizeof(double));
            if (r.U) std::memcpy(&uu[b * L2], r.U, L2 * sizeof(double));
            else
                for (int a = 0; a < L; ++a) uu[b * L2 + a * L + a] = 1.0;
        }
        DBuf<double> dh(B * L2), dV(B * L4), dU(B * L2), tmp(B * L4);
        hbufs.emplace_back(B * L2);
        vbufs.emplace_back(B * L4);
        CF_HIP(hipMemcpyAsync(dh.p, hh.data(), dh.bytes(), hipMemcpyHostToDevice, st));
        CF_HIP(hipMemcpyAsync(dV.p, vv.data(), dV.bytes(), hipMemcpyHostToDevice, st));
        CF_HIP(hipMemcpyAsync(dU.p, uu.data(), dU.bytes(), hipMemcpyHostToDevice, st));
        rotate_integrals_dev(L, B, dh.p, dV.p, dU.p, hbufs.back().p, vbufs.back().p, tmp.p, st);
        CF_HIP(hipStreamSynchronize(st));  // the host staging arrays go out of scope
        for (int b = 0; b < B; ++b) {
            hptr[g.second[b]] = hbufs.back().p + b * L2;
            vptr[g.second[b]] = vbufs.back().p + b * L4;
        }
    }
    // shared tables of each (L, pairs) shape, built once per context
    for (int i = 0; i < count; ++i) {
        if (!ok[i]) continue;
        const int L = reqs[i].L, np = reqs[i].nup;
        auto key = std::make_pair(L, np);
        if (ctx->doci_tabs.count(key)) continue;
        cf_ctx::DociTables& T = ctx->doci_tabs[key];
        const uint64_t n = binom_host(L, np);
        T.s.alloc(n);
        T.col.alloc(n * (uint64_t)np * (L - np));
        launch_unrank(L, np, 0, n, ctx->binom.p, T.s.p, st);
        DociEllArgs a;
        a.n = n;
        a.L = L;
        a.np = np;
        a.nv = L - np;
        a.s = T.s.p;
        launch_doci_ell_fill(a, ctx->binom.p, T.col.p, st);
    }
    // jobs by size class: one launch each
    uint64_t scratch_n = 0;
    for (int i = 0; i < count; ++i)
        if (ok[i]) scratch_n += 3 * results[i].n;
    DBuf<double> scratch(std::max<uint64_t>(scratch_n, 1)), dout(4 * (size_t)count);
    std::vector<DociBatchJob> jobs;
    std::vector<int> order;
    uint64_t off = 0;
    const int bounds[3] = {1024, 4096, cap};
    std::vector<std::pair<size_t, size_t>> ranges;  // [first, last) of each class in jobs
    int lo_n = 0;
    for (int c = 0; c < 3; ++c) {
        const size_t first = jobs.size();
        for (int i = 0; i < count; ++i) {
            if (!ok[i]) continue;
            const int n = (int)results[i].n;
            if (n <= lo_n || n > bounds[c]) continue;
            const cf_request& r = reqs[i];
            DociBatchJob J;
            J.h = hptr[i];
            J.V = vptr[i];
            const cf_ctx::DociTables& T = ctx->doci_tabs[std::make_pair(r.L, r.nup)];
            J.s = T.s.p;
            J.col = T.col.p;
            J.tol = r.tol > 0 ? r.tol : dflt.tol;
            J.tol_energy = r.tol_energy > 0 ? r.tol_energy : 0.0;
            J.seed = dflt.seed;
            J.out = dout.p + 4 * (size_t)i;
            J.scratch = scratch.p + off;
            off += 3 * (uint64_t)n;
            J.L = r.L;
            J.np = r.nup;
            J.nv = r.L - r.nup;
            J.n = n;
            J.check = r.check_every > 0 ? r.check_every : 2;
            J.max_steps = kBatchMaxSteps;
            jobs.push_back(J);
            order.push_back(i);
        }
        ranges.emplace_back(first, jobs.size());
        lo_n = bounds[c];
    }
    DBuf<DociBatchJob> djobs(std::max<size_t>(jobs.size(), 1));
    if (!jobs.empty())
        CF_HIP(hipMemcpyAsync(djobs.p, jobs.data(), jobs.size() * sizeof(DociBatchJob), hipMemcpyHostToDevice, st));
    for (int c = 0; c < 3; ++c) {
        const size_t a = ranges[c].first, b = ranges[c].second;
        if (b > a) launch_doci_batch_lanczos(djobs.p + a, (int)(b - a), bounds[c], st);
    }
    std::vector<double> hout(4 * (size_t)count, 0.0);
    CF_HIP(hipMemcpyAsync(hout.data(), dout.p, hout.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    CF_HIP(hipStreamSynchronize(st));
    for (int i : order) {
        cf_result& o = results[i];
        o.status = CF_OK;
        o.eigval = hout[4 * i + 0];
        o.energy = o.eigval + reqs[i].econst;
        o.resid = hout[4 * i + 1];
        o.n_matvec = (int)hout[4 * i + 2];
        o.n_matvec_f32 = 0;
        o.solve_status = (int)hout[4 * i + 3];
    }
    API_END
}

int cf_rdm_dense(cf_space* sp, const double* cvec, double* gamma, double* Gamma) {
    API_BEGIN
    CF_REQUIRE(sp && cvec && gamma, "NULL argument");
    cf_ctx* c = sp->ham->ctx;
    c->use();
    const int L = sp->ham->L;
    KTimer t(c, "rdm");
    if (sp->kind == CF_DETS)
        rdm_dets(c->stream, L, sp->n, sp->dup.p, sp->ddown.p, cvec, gamma, Gamma);
    else
        rdm_dense(c->stream, sp->kind, L, sp->ham->nup, sp->ham->ndown, sp->da, sp->db, sp->ta->s.p, sp->tb->s.p,
                  sp->ta->sub.p, sp->tb->sub.p, sp->ta->S, sp->tb->S, c->binom.p, cvec, sp->n, gamma, Gamma);
    API_END
}

}  // extern "C"
