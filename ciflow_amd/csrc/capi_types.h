// Handle types of the C ABI (include/ciflow_amd.h), shared by the files that
// implement it (capi.cpp, solve_batch.cpp).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include <rocblas/rocblas.h>

#include "common.h"
#include "kernels.h"
#include "sigma.h"
#include "spmv_doci.h"

using cf::DBuf;
using cf::Error;
using cf::ERR_HIP;
using cf::SigmaArgs;
using cf::SigmaShape;
using cf::str_t;
using cf::tls_alloc_stream;
using cf::HamDev;
using cf::Exc;

struct TimerFamily {
    int64_t launches = 0;
    double ms = 0.0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

struct cf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DBuf<uint64_t> binom;  // C(k, m): (kMaxL+1) x kBinomCols
    bool timing = false;
    std::map<std::string, TimerFamily> timers;
    // strings and hop columns of DOCI spaces, per (L, pairs), for batched solves
    struct DociTables {
        DBuf<str_t> s;
        DBuf<uint32_t> col;
    };
    std::map<std::pair<int, int>, DociTables> doci_tabs;
    // block tables of the block-staged DOCI product, per (L, pairs), for batched Davidson solves
    struct DociBlockTables {
        cf::DBuf<cf::DociBlock> blocks;
        cf::DBuf<uint16_t> rank_lo, lo_of;
        cf::DBuf<uint32_t> lo_off;
        cf::DociBlockArgs args;
    };
    std::map<std::pair<int, int>, DociBlockTables> doci_blk;
    // page-locked staging of batched Davidson solves (grown on demand)
    void* bpinned = nullptr;
    size_t bpinned_bytes = 0;
    void* batch_pinned(size_t bytes) {
        if (bytes > bpinned_bytes) {
            if (bpinned) CF_HIP(hipHostFree(bpinned));
            bpinned = nullptr;
            bpinned_bytes = 0;
            CF_HIP(hipHostMalloc(&bpinned, bytes, hipHostMallocDefault));
            bpinned_bytes = bytes;
        }
        return bpinned;
    }
    std::vector<hipEvent_t> pool;
    DBuf<double> hx, hy;  // staging for host-pointer products
    double* pinned = nullptr;  // kPinned doubles of page-locked memory for small read-backs
    // [0, 256): scalar read-backs; [256, 256 + 64*64): Ritz-vector coefficients
    static constexpr int kPinned = 256 + 64 * 64;
    rocblas_handle blas_ = nullptr;  // created on first use, bound to the stream

    rocblas_handle blas() {
        if (!blas_) {
            if (rocblas_create_handle(&blas_) != rocblas_status_success) {
                blas_ = nullptr;
                throw Error(ERR_HIP, "rocblas_create_handle failed");
            }
            if (rocblas_set_stream(blas_, stream) != rocblas_status_success)
                throw Error(ERR_HIP, "rocblas_set_stream failed");
        }
        return blas_;
    }

    void use() {
        CF_HIP(hipSetDevice(device));
        tls_alloc_stream = stream;
    }
    hipEvent_t event() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        CF_HIP(hipEventCreate(&e));
        return e;
    }
    void flush(TimerFamily& f) {
        for (auto& pr : f.pending) {
            CF_HIP(hipEventSynchronize(pr.second));
            float ms = 0.f;
            CF_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
            f.ms += ms;
            f.launches++;
            pool.push_back(pr.first);
            pool.push_back(pr.second);
        }
        f.pending.clear();
    }
    ~cf_ctx() {
        for (auto& kv : timers)
            for (auto& pr : kv.second.pending) {
                (void)hipEventDestroy(pr.first);
                (void)hipEventDestroy(pr.second);
            }
        for (auto e : pool) (void)hipEventDestroy(e);
        if (pinned) (void)hipHostFree(pinned);
        if (blas_) (void)rocblas_destroy_handle(blas_);
        // buffers ordered on the stream go back to the pool before it is destroyed
        binom.release();
        hx.release();
        hy.release();
        doci_tabs.clear();
        doci_blk.clear();
        if (bpinned) (void)hipHostFree(bpinned);
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (tls_alloc_stream == stream) tls_alloc_stream = nullptr;
    }
};

struct cf_ham {
    cf_ctx* ctx = nullptr;
    int L = 0, nup = 0, ndown = 0;
    double econst = 0.0;
    DBuf<double> dh, dV, dW, dJm;
    HamDev dev() const { return HamDev{L, dh.p, dV.p, dW.p, dJm.p}; }
};

struct StrTab {
    int L = 0, nel = 0;
    uint64_t d = 0;
    DBuf<str_t> s;
    DBuf<double> d1;
    int E = 0, S = 0;
    DBuf<Exc> full, sub;
};

struct cf_space {
    cf_ham* ham = nullptr;
    int kind = CF_FCI;
    uint64_t da = 0, db = 0, n = 0;
    std::unique_ptr<StrTab> ta, tb_own;
    StrTab* tb = nullptr;
    DBuf<uint32_t> segoff;
    DBuf<double> diag;
    DBuf<double> vec;  // resident eigenvector after cf_solve
    bool vec_valid = false;
    bool vec_sign_pending = false;  // sign convention applied when the vector is read
    uint64_t version = 0;  // incremented by cf_space_set_ham
    DBuf<str_t> dup, ddown;  // CF_DETS: the determinant list, list order
    uint64_t ascending = 0;  // CF_DETS: neighbours (i - 1, i) in ascending (up, down) order
};

struct cf_sigma {
    cf_space* sp = nullptr;
    bool f32 = false;
    DBuf<double> Ha, Hb_own;  // dense same-spin string Hamiltonians
    const double* Hb = nullptr;
    DBuf<float> x32;          // FP32 operand of the alpha-beta kernel
    rocblas_handle blas = nullptr;  // the context's handle
    SigmaArgs args;
    DBuf<double> vpair64, vself64;  // slice tables of the prefetching alpha-beta kernel
    DBuf<float> vpair32, vself32;
    SigmaShape shape;
    uint64_t version = 0;  // space Hamiltonian version the tables were made for
    // determinant-list spaces (CI_Big): the product on the fly from the
    // sorted keys, plus the stored CRS of the leading `stored` x `stored`
    // block (a space of the first `stored` determinants; both owned)
    bool dets = false;
    uint64_t stored = 0;
    DBuf<uint64_t> skey;
    DBuf<uint32_t> sidx;
    cf_space* block_sp = nullptr;
    cf_crs* block = nullptr;
};

namespace cf {
// One-string diagonal parts and excitation lists of a string table for H.
void fill_strtab(cf_ctx* ctx, const cf_ham* H, StrTab* t);
// Diagonal of the space for its current Hamiltonian.
void fill_space_diag(cf_space* sp);
// Strings of nel electrons in L orbitals with their H-dependent tables
// (excitation lists when `lists`), ordered on the context stream.
StrTab* make_strtab(cf_ctx* ctx, const cf_ham* H, int nel, bool lists);
// Block tables of the block-staged DOCI product for (L, np) with w low
// orbitals; T.args gets everything but diag and zero.
void make_doci_block_tables(cf_ctx* c, int L, int np, int w, cf_ctx::DociBlockTables& T);
}  // namespace cf
