// The bindings a CIFlow maintainer would add, compiled with g++
// against include/ciflow_amd.h and linked to libciflow_amd.so:
//   Route 1: a matrix class overriding the virtual mvprod(const double*, double*)
//            seam (SparseMatrix_CRS.h:92,126) that forwards to cf_mvprod;
//   Route 3: the dsaupd_-shaped reverse-communication loop of
//            SparseMatrix_CRS.cpp:441-449 driving that class through the seam;
//   Route 2: cf_solve, the GPU-resident arpackDiagonalize;
// on FCI L=8 4a4b with seeded integrals (seed 0, 4,900 determinants), checked
// against each other and against the dense-eigh energy -12.512862152710419.
// The base class below stands in for SparseMatrix_CRS's virtual interface
// (the reference's own header is not part of this repository).
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "ciflow_amd.h"

static void check(int st, const char* what) {
    if (st != CF_OK) throw std::runtime_error(std::string(what) + ": " + cf_last_error());
}

// the seam: SparseMatrix_CRS::mvprod is virtual and arpackDiagonalize calls it
class MatVecSeam {
  public:
    explicit MatVecSeam(unsigned n) : n_(n) {}
    virtual ~MatVecSeam() = default;
    virtual void mvprod(const double* x, double* y) const = 0;
    unsigned getn() const { return n_; }

  private:
    unsigned n_;
};

// Route 1: the GPU matrix behind the seam
class SparseMatrix_CRS_Device : public MatVecSeam {
  public:
    SparseMatrix_CRS_Device(cf_crs* m, unsigned n) : MatVecSeam(n), m_(m) {}
    ~SparseMatrix_CRS_Device() override { cf_crs_destroy(m_); }
    void mvprod(const double* x, double* y) const override {
        if (cf_mvprod(m_, x, y) != CF_OK) throw std::runtime_error(cf_last_error());
    }
    cf_crs* handle() const { return m_; }

  private:
    cf_crs* m_;
};

// Route 3: the reverse-communication loop, unchanged in shape
static double rci_lowest(const MatVecSeam& mat, std::vector<double>& eigvec, int& info) {
    const unsigned n = mat.getn();
    cf_rci* r = nullptr;
    int ido = 0, ipntr[3] = {0, 0, 0};
    check(cf_rci_create(n, 42, 1e-12, 0, 20240601, &r), "cf_rci_create");
    std::vector<double> workd(3 * (size_t)n);
    for (;;) {
        check(cf_rci_step(r, &ido, workd.data(), ipntr), "cf_rci_step");
        if (ido == 99) break;
        mat.mvprod(&workd[ipntr[0] - 1], &workd[ipntr[1] - 1]);
    }
    double ev = 0.0;
    eigvec.assign(n, 0.0);
    check(cf_rci_result(r, &ev, eigvec.data(), &info), "cf_rci_result");
    cf_rci_destroy(r);
    return ev;
}

int main() {
    const int L = 8, nup = 4, ndown = 4;
    const double E_ref = -12.512862152710419;  // dense eigh of the oracle's matrix
    int fails = 0;
    try {
        cf_ctx* ctx = nullptr;
        cf_ham* ham = nullptr;
        cf_space* sp = nullptr;
        cf_crs* m = nullptr;
        check(cf_ctx_create(0, &ctx), "cf_ctx_create");
        std::vector<double> h(L * L), V((size_t)L * L * L * L);
        check(cf_gen_integrals(L, 0, h.data(), V.data()), "cf_gen_integrals");
        check(cf_ham_create(ctx, L, nup, ndown, h.data(), V.data(), 0.0, &ham), "cf_ham_create");
        check(cf_space_create(ham, CF_FCI, &sp), "cf_space_create");
        check(cf_crs_build(sp, CF_F64, 1e-14, &m), "cf_crs_build");
        uint64_t n = 0, da = 0, db = 0;
        check(cf_space_dims(sp, &n, &da, &db), "cf_space_dims");
        {
            SparseMatrix_CRS_Device mat(m, (unsigned)n);
            std::vector<double> v1;
            int info = -1;
            const double e1 = rci_lowest(mat, v1, info);
            std::printf("route 1+3 (RCI loop, GPU mvprod through the seam): E = %.12f info %d\n", e1, info);
            if (info != 0 || std::fabs(e1 - E_ref) > 1e-10) ++fails;

            cf_solve_opts o;
            cf_solve_opts_default(&o);
            cf_solve_info si;
            std::vector<double> v2(n);
            check(cf_solve(mat.handle(), &o, v2.data(), &si), "cf_solve");
            std::printf("route 2 (cf_solve): E = %.12f, %d products, status %d\n", si.energy, si.n_matvec, si.status);
            if (si.status != 0 || std::fabs(si.energy - E_ref) > 1e-10) ++fails;
            // the two eigenvectors agree up to sign (both use the largest-component-positive convention)
            double d = 0.0;
            for (uint64_t i = 0; i < n; ++i) d = std::fmax(d, std::fabs(v1[i] - v2[i]));
            std::printf("eigenvector difference %.2e\n", d);
            if (d > 1e-6) ++fails;
            // CIDens layout: one[2][L(L+1)/2]; tr gamma = N
            std::vector<double> one(L * (L + 1)), two(3 * (size_t)L * L * (L * L + 1) / 2);
            check(cf_rdm(sp, v2.data(), one.data(), two.data()), "cf_rdm");
            double tr = 0.0;
            for (int s = 0; s < 2; ++s)
                for (int p = 0; p < L; ++p) tr += one[s * (L * (L + 1) / 2) + p * L + p - p * (p + 1) / 2];
            std::printf("tr gamma = %.12f\n", tr);
            if (std::fabs(tr - (nup + ndown)) > 1e-10) ++fails;
            // the reference's 32-bit row pointer
            uint64_t cnt = 0;
            check(cf_crs_count_upper(mat.handle(), 0, n, &cnt), "cf_crs_count_upper");
            std::vector<uint32_t> row32(n + 1), col(cnt);
            std::vector<double> val(cnt);
            check(cf_crs_export_upper_u32(mat.handle(), 0, n, row32.data(), col.data(), val.data()),
                  "cf_crs_export_upper_u32");
            std::printf("upper nonzeros %llu (row[n] = %u)\n", (unsigned long long)cnt, row32[n]);
            if (cnt != 886900 || row32[n] != cnt) ++fails;
        }  // the device matrix releases its cf_crs
        check(cf_space_destroy(sp), "cf_space_destroy");
        check(cf_ham_destroy(ham), "cf_ham_destroy");
        check(cf_ctx_destroy(ctx), "cf_ctx_destroy");
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 2;
    }
    if (fails) {
        std::printf("%d check(s) failed\n", fails);
        return 1;
    }
    std::printf("routes ok\n");
    return 0;
}
