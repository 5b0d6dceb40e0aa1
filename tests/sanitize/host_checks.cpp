// Checks of the host-only C++ under AddressSanitizer and UndefinedBehaviorSanitizer:
// the library's file parsers, seeded generators, exp(K) and reverse-communication
// Lanczos (host.cpp, integrals.cpp, lanczos.cpp) and the CPU oracle
// (oracle/ciflow_oracle.cpp), built together with g++ by tests/test_sanitizers.py.
// Usage: host_checks <tests/golden directory> <scratch directory>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../include/ciflow_amd.h"

extern "C" {
void or_gen_integrals(int L, uint64_t seed, double* h, double* V);
uint64_t or_binom(int n, int k);
void* or_build_crs(int kind, int L, int nup, int ndown, const double* h, const double* V, double zero, int start_l,
                   int end_l, int num_t);
uint64_t or_crs_nrows(void* m);
uint64_t or_crs_nnz(void* m);
void or_crs_copy(void* m, uint64_t* row, uint32_t* col, double* val);
void or_crs_free(void* m);
void* or_crs_full_from_upper(uint64_t n, const uint64_t* row, const uint32_t* col, const double* data);
void or_mvprod_sym(uint64_t n, const uint64_t* row, const uint32_t* col, const double* data, const double* x,
                   double* y);
void or_mvprod_full(uint64_t n, const uint64_t* row, const uint32_t* col, const double* data, const double* x,
                    double* y, int num_t);
int or_lanczos(uint64_t n, const uint64_t* row, const uint32_t* col, const double* data, int ncv, double tol,
               int maxrestart, int num_t, double* eigval, double* eigvec, int* n_matvec, int* n_restart);
void or_rdm_fci(int L, int nup, int ndown, const double* c, double* one, double* two);
void or_rdm_doci(int L, int npair, const double* c, double* one, double* two);
double or_dens_energy(int L, const double* h, const double* V, double econst, const double* one, const double* two);
void or_rotate_integrals(int L, const double* U, const double* h, const double* V, double* ho, double* Vo);
int64_t or_read_dets(const char* path, uint64_t cap, uint64_t* up, uint64_t* down);
void* or_build_crs_dets(int L, int nup, int ndown, const double* h, const double* V, double zero, uint64_t n,
                        const uint64_t* up, const uint64_t* down, int num_t);
double or_ham_element(int L, int nup, int ndown, const double* h, const double* V, uint64_t up1, uint64_t down1,
                      uint64_t up2, uint64_t down2);
}

static int failures = 0;
#define CHECK(cond, ...)                                    \
    do {                                                    \
        if (!(cond)) {                                      \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                       \
            std::printf("\n");                              \
            ++failures;                                     \
        }                                                   \
    } while (0)

struct Crs {
    std::vector<uint64_t> row;
    std::vector<uint32_t> col;
    std::vector<double> val;
    uint64_t n() const { return row.size() - 1; }
};

static Crs take(void* m) {
    Crs c;
    c.row.resize(or_crs_nrows(m) + 1);
    c.col.resize(or_crs_nnz(m));
    c.val.resize(or_crs_nnz(m));
    or_crs_copy(m, c.row.data(), c.col.data(), c.val.data());
    or_crs_free(m);
    return c;
}

// Lowest eigenvalue of a small dense symmetric matrix by cyclic Jacobi.
static double jacobi_lowest(std::vector<double> A, int n) {
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
        if (off < 1e-26) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
            }
    }
    double m = A[0];
    for (int i = 1; i < n; ++i) m = std::min(m, A[i * n + i]);
    return m;
}

static void check_parsers(const std::string& gold, const std::string& tmp) {
    const std::string dump = gold + "/beh2_sto3g_1_5.dat";
    int L = 0, na = 0, nb = 0;
    double ec = 0.0;
    CHECK(cf_read_integrals(dump.c_str(), &L, &na, &nb, &ec, nullptr, nullptr, nullptr) == CF_OK, "size query");
    CHECK(L == 7 && na == 3 && nb == 3, "BeH2 header: L=%d na=%d nb=%d", L, na, nb);
    std::vector<double> h(L * L), V((size_t)L * L * L * L);
    std::vector<int> ir(L);
    CHECK(cf_read_integrals(dump.c_str(), &L, &na, &nb, &ec, h.data(), V.data(), ir.data()) == CF_OK, "read");
    double asym = 0.0;
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) {
            asym = std::max(asym, std::fabs(h[i * L + j] - h[j * L + i]));
            for (int k = 0; k < L; ++k)
                for (int l = 0; l < L; ++l) {
                    const double v = V[((i * L + j) * L + k) * L + l];
                    asym = std::max(asym, std::fabs(v - V[((j * L + i) * L + l) * L + k]));
                    asym = std::max(asym, std::fabs(v - V[((k * L + l) * L + i) * L + j]));
                }
        }
    CHECK(asym == 0.0, "integral symmetry %g", asym);
    // malformed files: each must fail with a status and a message, not crash
    const char* bad[] = {
        "no header at all\n",
        "****  Molecular Integrals\nNalpha = 1\nNbeta = 1\nSymmetry Label = c1\nNirreps = 1\n"
        "Nuclear Repulsion Energy = 0.5\nNumber Of Molecular Orbitals = 2\nIrreps\n0 0\n"
        "****  One-electron\n0 5 1.0\n",
        "****  Molecular Integrals\nNalpha = 1\nNbeta = 1\nSymmetry Label = c1\nNirreps = 1\n"
        "Nuclear Repulsion Energy = 0.5\nNumber Of Molecular Orbitals = 2\nIrreps\n0 0\n"
        "****  One-electron\n0 0 1.0\n****  Two\n0 0 0 9 1.0\n",
        "****  Molecular Integrals\nNalpha = 1\nNbeta = 1\nSymmetry Label = c1\nNirreps = 1\n"
        "Nuclear Repulsion Energy = 0.5\nNumber Of Molecular Orbitals = 200\n",
    };
    for (size_t b = 0; b < sizeof(bad) / sizeof(bad[0]); ++b) {
        const std::string path = tmp + "/bad" + std::to_string(b) + ".dat";
        std::ofstream(path) << bad[b];
        std::vector<double> hh(4), VV(16);
        int st = cf_read_integrals(path.c_str(), &L, &na, &nb, &ec, nullptr, nullptr, nullptr);
        if (st == CF_OK) st = cf_read_integrals(path.c_str(), &L, &na, &nb, &ec, hh.data(), VV.data(), nullptr);
        CHECK(st != CF_OK && std::strlen(cf_last_error()) > 0, "malformed integral file %zu accepted", b);
    }
    CHECK(cf_read_integrals((tmp + "/missing.dat").c_str(), &L, &na, &nb, &ec, nullptr, nullptr, nullptr) ==
              CF_ERR_ARG, "missing file");
    // determinant files against the oracle's reader
    for (const char* f : {"fci_beh2_sto-3gdeterminants.dat", "GKCI_Be_cntSp_cc-pVDZ_0100.txt"}) {
        const std::string path = gold + "/" + f;
        uint64_t cnt = 0;
        CHECK(cf_read_determinants(path.c_str(), 0, nullptr, nullptr, &cnt) == CF_OK, "count %s", f);
        std::vector<uint64_t> up(cnt), dn(cnt), oup(cnt), odn(cnt);
        CHECK(cf_read_determinants(path.c_str(), cnt, up.data(), dn.data(), &cnt) == CF_OK, "read %s", f);
        const int64_t on = or_read_dets(path.c_str(), cnt, oup.data(), odn.data());
        CHECK(on == (int64_t)cnt && up == oup && dn == odn, "determinants of %s differ from the oracle", f);
        if (cnt > 3) {  // a short cap stores the first entries and still counts all
            std::vector<uint64_t> u3(3), d3(3);
            uint64_t c3 = 0;
            CHECK(cf_read_determinants(path.c_str(), 3, u3.data(), d3.data(), &c3) == CF_OK && c3 == cnt &&
                      u3[2] == up[2], "capped read of %s", f);
        }
    }
    const std::string longdet = tmp + "/long.dat";
    std::ofstream(longdet) << std::string(70, '1') << "|" << "1\n";
    uint64_t cnt = 0;
    CHECK(cf_read_determinants(longdet.c_str(), 0, nullptr, nullptr, &cnt) == CF_ERR_UNSUPPORTED, "65+ orbitals");
}

static void check_generators() {
    for (int L = 1; L <= 16; ++L) {
        std::vector<double> h(L * L), V((size_t)L * L * L * L), oh(h.size()), oV(V.size());
        CHECK(cf_gen_integrals(L, 7 + L, h.data(), V.data()) == CF_OK, "gen L=%d", L);
        or_gen_integrals(L, 7 + L, oh.data(), oV.data());
        CHECK(h == oh && V == oV, "generator differs from the oracle at L=%d", L);
    }
    for (int L = 1; L <= 24; ++L) {
        std::vector<double> K(L * L), U(L * L);
        CHECK(cf_gen_rotation(L, 3 * L, 0.3, K.data()) == CF_OK, "rotation L=%d", L);
        CHECK(cf_expm_skew(L, K.data(), U.data()) == CF_OK, "expm L=%d", L);
        double err = 0.0;
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < L; ++j) {
                double s = 0.0;
                for (int k = 0; k < L; ++k) s += U[i * L + k] * U[j * L + k];
                err = std::max(err, std::fabs(s - (i == j ? 1.0 : 0.0)));
                if (i <= j) err = std::max(err, std::fabs(K[i * L + j] + K[j * L + i]));
            }
        CHECK(err < 1e-13, "exp(K) not orthogonal or K not antisymmetric at L=%d: %g", L, err);
    }
    CHECK(cf_gen_integrals(0, 1, nullptr, nullptr) == CF_ERR_ARG, "bad generator arguments accepted");
}

static void check_rci() {
    std::mt19937_64 rng(11);
    std::normal_distribution<double> nd;
    const int sizes[][2] = {{1, 2}, {2, 2}, {5, 3}, {37, 8}, {200, 20}, {300, 42}};
    for (const auto& sz : sizes) {
        const int n = sz[0], ncv = sz[1];
        std::vector<double> A((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) A[i * n + j] = A[j * n + i] = nd(rng) + (i == j ? 0.1 * i : 0.0);
        cf_rci* r = nullptr;
        CHECK(cf_rci_create(n, ncv, 1e-12, 0, 5, &r) == CF_OK, "rci create n=%d", n);
        std::vector<double> workd(3 * (size_t)n);
        int ido = 0, ipntr[3] = {0, 0, 0}, steps = 0;
        while (true) {
            CHECK(cf_rci_step(r, &ido, workd.data(), ipntr) == CF_OK, "rci step n=%d", n);
            if (ido != 1 || ++steps > 100000) break;
            const double* x = workd.data() + ipntr[0] - 1;
            double* y = workd.data() + ipntr[1] - 1;
            for (int i = 0; i < n; ++i) {
                double s = 0.0;
                for (int j = 0; j < n; ++j) s += A[i * n + j] * x[j];
                y[i] = s;
            }
        }
        double ev = 0.0;
        int info = -1;
        std::vector<double> vec(n);
        CHECK(cf_rci_result(r, &ev, vec.data(), &info) == CF_OK && info == 0, "rci result n=%d info=%d", n, info);
        const double ref = jacobi_lowest(A, n);
        CHECK(std::fabs(ev - ref) <= 1e-9 * std::max(1.0, std::fabs(ref)), "rci n=%d: %.15g vs %.15g", n, ev, ref);
        CHECK(cf_rci_step(r, &ido, workd.data(), ipntr) == CF_ERR_ARG, "step after the end accepted");
        cf_rci_destroy(r);
    }
    cf_rci* r = nullptr;
    CHECK(cf_rci_create(10, 1, 1e-12, 0, 1, &r) == CF_ERR_ARG, "ncv 1 accepted");
}

static void check_oracle() {
    const int L = 6, na = 3, nb = 2;
    std::vector<double> h(L * L), V((size_t)L * L * L * L);
    or_gen_integrals(L, 4, h.data(), V.data());
    const int lines = (int)or_binom(L, na);
    Crs up = take(or_build_crs(0, L, na, nb, h.data(), V.data(), 1e-14, 0, lines, 2));
    Crs full = take(or_crs_full_from_upper(up.n(), up.row.data(), up.col.data(), up.val.data()));
    const uint64_t n = up.n();
    CHECK(n == or_binom(L, na) * or_binom(L, nb), "FCI dimension");
    std::vector<double> x(n), y1(n), y2(n);
    for (uint64_t i = 0; i < n; ++i) x[i] = std::sin(0.1 * i + 1.0);
    or_mvprod_sym(n, up.row.data(), up.col.data(), up.val.data(), x.data(), y1.data());
    or_mvprod_full(n, full.row.data(), full.col.data(), full.val.data(), x.data(), y2.data(), 2);
    double d = 0.0;
    for (uint64_t i = 0; i < n; ++i) d = std::max(d, std::fabs(y1[i] - y2[i]));
    CHECK(d < 1e-12, "symmetric and full products differ by %g", d);
    // spot check: an element of the CRS against get_ham_element
    std::vector<double> dense(n * n, 0.0);
    for (uint64_t i = 0; i < n; ++i)
        for (uint64_t k = full.row[i]; k < full.row[i + 1]; ++k) dense[i * n + full.col[k]] = full.val[k];
    double E = 0.0;
    std::vector<double> c(n);
    int nmv = 0, nrs = 0;
    CHECK(or_lanczos(n, full.row.data(), full.col.data(), full.val.data(), 20, 1e-12, 1000, 2, &E, c.data(), &nmv,
                     &nrs) == 0, "oracle Lanczos");
    std::vector<double> Ad(dense);
    const double ref = jacobi_lowest(Ad, (int)n);
    CHECK(std::fabs(E - ref) < 1e-10, "oracle Lanczos %.15g vs Jacobi %.15g", E, ref);
    std::vector<double> one(L * (L + 1)), two(3 * (size_t)L * L * (L * L + 1) / 2);
    or_rdm_fci(L, na, nb, c.data(), one.data(), two.data());
    const double Ed = or_dens_energy(L, h.data(), V.data(), 0.0, one.data(), two.data());
    CHECK(std::fabs(Ed - E) < 1e-10, "density energy %.15g vs %.15g", Ed, E);
    // DOCI and its RDMs
    Crs dup = take(or_build_crs(1, L, 3, 3, h.data(), V.data(), 1e-14, 0, (int)or_binom(L, 3), 2));
    Crs dfull = take(or_crs_full_from_upper(dup.n(), dup.row.data(), dup.col.data(), dup.val.data()));
    std::vector<double> cd(dup.n());
    CHECK(or_lanczos(dup.n(), dfull.row.data(), dfull.col.data(), dfull.val.data(), 10, 1e-12, 1000, 1, &E, cd.data(),
                     &nmv, &nrs) == 0, "DOCI Lanczos");
    or_rdm_doci(L, 3, cd.data(), one.data(), two.data());
    CHECK(std::fabs(or_dens_energy(L, h.data(), V.data(), 0.0, one.data(), two.data()) - E) < 1e-10, "DOCI density");
    // rotation round trip
    std::vector<double> K(L * L), U(L * L), Ut(L * L), ho(h.size()), Vo(V.size()), hb(h.size()), Vb(V.size());
    cf_gen_rotation(L, 9, 0.4, K.data());
    cf_expm_skew(L, K.data(), U.data());
    for (int a = 0; a < L; ++a)
        for (int i = 0; i < L; ++i) Ut[a * L + i] = U[i * L + a];
    or_rotate_integrals(L, U.data(), h.data(), V.data(), ho.data(), Vo.data());
    or_rotate_integrals(L, Ut.data(), ho.data(), Vo.data(), hb.data(), Vb.data());
    double rt = 0.0;
    for (size_t i = 0; i < V.size(); ++i) rt = std::max(rt, std::fabs(Vb[i] - V[i]));
    CHECK(rt < 1e-13, "rotation round trip %g", rt);
    // determinant-list build on a shuffled subset of the FCI space: elements equal get_ham_element
    std::vector<uint64_t> upd, dnd;
    for (uint64_t a = 0; a < (1u << L); ++a)
        for (uint64_t b = 0; b < (1u << L); ++b)
            if (__builtin_popcountll(a) == na && __builtin_popcountll(b) == nb && ((a * 7 + b * 3) % 5 != 0)) {
                upd.push_back(a);
                dnd.push_back(b);
            }
    Crs dl = take(or_build_crs_dets(L, na, nb, h.data(), V.data(), 1e-14, upd.size(), upd.data(), dnd.data(), 2));
    double de = 0.0;
    for (uint64_t i = 0; i < dl.n(); i += 7)
        for (uint64_t k = dl.row[i]; k < dl.row[i + 1]; ++k)
            de = std::max(de, std::fabs(dl.val[k] - or_ham_element(L, na, nb, h.data(), V.data(), upd[i], dnd[i],
                                                                   upd[dl.col[k]], dnd[dl.col[k]])));
    CHECK(de < 1e-12, "determinant-list elements differ by %g", de);
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: host_checks <golden dir> <scratch dir>\n");
        return 2;
    }
    check_parsers(argv[1], argv[2]);
    check_generators();
    check_rci();
    check_oracle();
    if (failures) {
        std::printf("%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("all host checks passed\n");
    return 0;
}
