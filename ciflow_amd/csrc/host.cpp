// Host-only entry points of the C ABI: error text, the integral and
// determinant file parsers, the seeded generators and exp(K), and the
// reverse-communication Lanczos (cf_rci_*).  No HIP: this file builds with a
// plain C++ compiler (the CPU sanitizer build links it with lanczos.cpp and
// integrals.cpp).
#include "../../include/ciflow_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "errors.h"
#include "host.h"
#include "lanczos.h"

namespace cf {
void gen_integrals(int L, uint64_t seed, double* h, double* V);
void gen_rotation(int L, uint64_t seed, double sigma, double* K);
void expm_skew(int L, const double* K, double* U);

namespace {
thread_local std::string g_err;
}  // namespace

int set_last_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace cf

using namespace cf;

#define API_BEGIN try {
#define API_END                                                        \
    }                                                                  \
    catch (const cf::Error& e) {                                       \
        return set_last_error(e.code, e.what());                       \
    }                                                                  \
    catch (const std::bad_alloc&) {                                    \
        return set_last_error(CF_ERR_NOMEM, "host allocation failed"); \
    }                                                                  \
    catch (const std::exception& e) {                                  \
        return set_last_error(CF_ERR_HIP, e.what());                   \
    }                                                                  \
    return CF_OK;

namespace cf {
void sign_fix(std::vector<double>& x) {
    size_t im = 0;
    for (size_t q = 1; q < x.size(); ++q)
        if (std::fabs(x[q]) > std::fabs(x[im])) im = q;
    if (!x.empty() && x[im] < 0)
        for (auto& v : x) v = -v;
}

}  // namespace cf

namespace {
// Krylov basis in host memory for the reverse-communication interface (the
// caller applies the operator to host vectors).
struct HostBackend : KrylovBackend {
    uint64_t n;
    int M;
    std::vector<double> Vb, w;
    HostBackend(uint64_t n_, int ncv) : n(n_) {
        M = (uint64_t)ncv > n ? (int)n : ncv;
        Vb.assign((size_t)(M + 1) * n, 0.0);
        w.assign(n, 0.0);
    }
    double* v(int i) { return Vb.data() + (size_t)i * n; }
    double dot(const double* a, const double* b) {
        double s = 0.0;
        for (uint64_t q = 0; q < n; ++q) s += a[q] * b[q];
        return s;
    }
    void start_random(uint64_t seed) override {
        double* v0 = v(0);
        for (uint64_t q = 0; q < n; ++q) {
            uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (q + 1);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            z ^= z >> 31;
            v0[q] = 1.0 + 0.5 * ((double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5);
        }
        normalize0();
    }
    void start_keep() override { normalize0(); }
    void apply(int) override { throw Error(ERR_STATE, "host backend is driven by reverse communication"); }
    void orth(int j, double* alpha, double* beta) override {
        double a = 0.0;
        std::vector<double> c(j + 1);
        for (int pass = 0; pass < 2; ++pass) {
            for (int i = 0; i <= j; ++i) c[i] = dot(v(i), w.data());
            a += c[j];
            for (int i = 0; i <= j; ++i) {
                const double* vi = v(i);
                for (uint64_t q = 0; q < n; ++q) w[q] -= c[i] * vi[q];
            }
        }
        *alpha = a;
        *beta = std::sqrt(dot(w.data(), w.data()));
    }
    void set_next(int j1, double beta) override {
        double* d = v(j1);
        for (uint64_t q = 0; q < n; ++q) d[q] = w[q] / beta;
    }
    void rotate(int m, int k, const double* Y) override {
        std::vector<double> row(m);
        for (uint64_t q = 0; q < n; ++q) {
            for (int i = 0; i < m; ++i) row[i] = Vb[(size_t)i * n + q];
            for (int c = 0; c < k; ++c) {
                double s = 0.0;
                for (int i = 0; i < m; ++i) s += row[i] * Y[(size_t)i * m + c];
                Vb[(size_t)c * n + q] = s;
            }
        }
    }
    void move(int from, int to) override { std::memcpy(v(to), v(from), n * sizeof(double)); }
    void normalize0() override {
        double s = 1.0 / std::sqrt(dot(v(0), v(0)));
        double* v0 = v(0);
        for (uint64_t q = 0; q < n; ++q) v0[q] *= s;
    }
    void rayleigh(double* theta, double* resid) override {
        double th = dot(v(0), w.data());
        double r2 = 0.0;
        const double* v0 = v(0);
        for (uint64_t q = 0; q < n; ++q) {
            double r = w[q] - th * v0[q];
            r2 += r * r;
        }
        *theta = th;
        *resid = std::sqrt(r2);
    }
};
}  // namespace

struct cf_rci {
    std::unique_ptr<HostBackend> b;
    std::unique_ptr<ThickRestartLanczos> sm;
    uint64_t n = 0;
    bool started = false, done = false;
};

extern "C" {

int cf_version(void) { return 100; }

const char* cf_last_error(void) { return g_err.c_str(); }

int cf_gen_integrals(int L, uint64_t seed, double* h, double* V) {
    API_BEGIN
    CF_REQUIRE(L >= 1 && L <= kMaxL && h && V, "bad arguments");
    gen_integrals(L, seed, h, V);
    API_END
}

int cf_expm_skew(int L, const double* K, double* U) {
    API_BEGIN
    CF_REQUIRE(L >= 1 && L <= kMaxL && K && U, "bad arguments");
    expm_skew(L, K, U);
    API_END
}

int cf_gen_rotation(int L, uint64_t seed, double sigma, double* K) {
    API_BEGIN
    CF_REQUIRE(L >= 1 && L <= kMaxL && K, "bad arguments");
    gen_rotation(L, seed, sigma, K);
    API_END
}

int cf_read_integrals(const char* path, int* L, int* nup, int* ndown, double* econst, double* h, double* V,
                      int* irreps) {
    API_BEGIN
    CF_REQUIRE(path && L && nup && ndown && econst, "NULL argument");
    std::ifstream f(path);
    if (!f.is_open()) throw Error(ERR_ARG, std::string("cannot open ") + path);
    std::string line;
    auto next = [&](const char* what) {
        if (!std::getline(f, line)) throw Error(ERR_ARG, std::string("integral file ends before ") + what);
        while (!line.empty() && line.back() == '\r') line.pop_back();
    };
    auto value_of = [&](const char* what) -> std::string {
        next(what);
        const size_t eq = line.find('=');
        if (eq == std::string::npos) throw Error(ERR_ARG, std::string("integral file: no '=' on the ") + what + " line");
        return line.substr(eq + 1);
    };
    do next("the \"****  Molecular Integrals\" header");
    while (line.find("****  Molecular Integrals") == std::string::npos);
    const int na = std::atoi(value_of("Nalpha").c_str());
    const int nb = std::atoi(value_of("Nbeta").c_str());
    value_of("Symmetry Label");
    value_of("Nirreps");
    const double ec = std::strtod(value_of("Nuclear Repulsion Energy").c_str(), nullptr);
    const int nL = std::atoi(value_of("Number Of Molecular Orbitals").c_str());
    CF_REQUIRE(nL >= 1 && nL <= kMaxL, "integral file: number of orbitals out of range");
    next("the irrep header");
    next("the orbital irreps");
    {
        std::istringstream is(line);
        for (int i = 0; i < nL; ++i) {
            int ir = 0;
            if (!(is >> ir)) throw Error(ERR_ARG, "integral file: fewer orbital irreps than orbitals");
            if (irreps) irreps[i] = ir;
        }
    }
    *L = nL;
    *nup = na;
    *ndown = nb;
    *econst = ec;
    if (!h && !V) return CF_OK;
    CF_REQUIRE(h && V, "h and V must both be given");
    const size_t L2 = (size_t)nL * nL, L4 = L2 * L2;
    std::fill(h, h + L2, 0.0);
    std::fill(V, V + L4, 0.0);
    do next("the one-electron integrals");
    while (line.rfind("**", 0) != 0);  // lines up to the first starting with "**"
    auto bad = [&](int x) { return x < 0 || x >= nL; };
    for (;;) {  // one-electron lines up to the next line starting with '*'
        next("the two-electron integrals");
        if (line.empty()) continue;
        if (line[0] == '*') break;
        std::istringstream is(line);
        int i, j;
        double v;
        if (!(is >> i >> j >> v) || bad(i) || bad(j)) throw Error(ERR_ARG, "integral file: bad one-electron line: " + line);
        h[(size_t)i * nL + j] = v;
        h[(size_t)j * nL + i] = v;
    }
    auto at = [&](int p, int q, int r, int t) -> double& { return V[(((size_t)p * nL + q) * nL + r) * nL + t]; };
    for (;;) {  // two-electron lines up to the next line starting with '*'
        next("the end of the two-electron integrals");
        if (line.empty()) continue;
        if (line[0] == '*') break;
        std::istringstream is(line);
        int a, b, c, d;
        double v;
        if (!(is >> a >> b >> c >> d >> v) || bad(a) || bad(b) || bad(c) || bad(d))
            throw Error(ERR_ARG, "integral file: bad two-electron line: " + line);
        // chemist (ab|cd) and its 8-fold symmetry, stored as physics <ac|bd>
        const int P[8][4] = {{a, b, c, d}, {b, a, c, d}, {a, b, d, c}, {b, a, d, c},
                             {c, d, a, b}, {d, c, a, b}, {c, d, b, a}, {d, c, b, a}};
        for (const auto& x : P) at(x[0], x[2], x[1], x[3]) = v;
    }
    API_END
}

int cf_read_determinants(const char* path, uint64_t cap, uint64_t* up, uint64_t* down, uint64_t* count) {
    API_BEGIN
    CF_REQUIRE(path && count, "NULL argument");
    CF_REQUIRE(cap == 0 || (up && down), "NULL output arrays");
    std::ifstream f(path);
    if (!f.is_open()) throw Error(ERR_ARG, std::string("cannot open ") + path);
    std::string line;
    uint64_t n = 0;
    while (std::getline(f, line)) {
        while (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.find('#') != std::string::npos) continue;
        const size_t cut = line.find('|');
        if (cut == std::string::npos) continue;
        if (cut > 64)
            throw Error(ERR_UNSUPPORTED, "determinant strings longer than 64 orbitals");
        const uint64_t u = std::strtoull(line.substr(0, cut).c_str(), nullptr, 2);
        const uint64_t d = std::strtoull(line.substr(cut + 1).c_str(), nullptr, 2);
        if (n < cap) {
            up[n] = u;
            down[n] = d;
        }
        ++n;
    }
    *count = n;
    API_END
}

int cf_rci_create(uint64_t n, int ncv, double tol, int maxrestart, uint64_t seed, cf_rci** out) {
    API_BEGIN
    CF_REQUIRE(out && n > 0, "bad arguments");
    CF_REQUIRE(ncv >= 2 && ncv <= 63 && tol > 0.0, "bad ncv / tol");
    std::unique_ptr<cf_rci> r(new cf_rci());
    r->n = n;
    r->b.reset(new HostBackend(n, ncv));
    int maxr = maxrestart > 0 ? maxrestart : (int)std::min<uint64_t>(3 * n, 1u << 30);
    r->sm.reset(new ThickRestartLanczos(r->b.get(), n, ncv, tol, maxr, seed, false));
    *out = r.release();
    API_END
}

int cf_rci_step(cf_rci* r, int* ido, double* workd, int* ipntr) {
    API_BEGIN
    CF_REQUIRE(r && ido && workd && ipntr, "NULL argument");
    CF_REQUIRE(!r->done, "reverse communication already finished");
    const uint64_t n = r->n;
    if (*ido == 0) {
        CF_REQUIRE(!r->started, "ido == 0 after the iteration started");
        r->started = true;
    } else if (*ido == 1) {
        CF_REQUIRE(r->started, "ido == 1 before the first call");
        std::memcpy(r->b->w.data(), workd + n, n * sizeof(double));
    } else {
        throw Error(ERR_ARG, "ido must be 0 or 1");
    }
    if (r->sm->step() == ThickRestartLanczos::NEED_APPLY) {
        std::memcpy(workd, r->b->v(r->sm->pending()), n * sizeof(double));
        ipntr[0] = 1;
        ipntr[1] = (int)n + 1;
        *ido = 1;
    } else {
        r->done = true;
        *ido = 99;
    }
    API_END
}

int cf_rci_result(cf_rci* r, double* eigval, double* eigvec, int* info) {
    API_BEGIN
    CF_REQUIRE(r && r->done, "reverse communication not finished");
    if (eigval) *eigval = r->sm->eigval();
    if (info) *info = r->sm->status();
    if (eigvec) {
        std::vector<double> x(r->b->v(0), r->b->v(0) + r->n);
        sign_fix(x);
        std::memcpy(eigvec, x.data(), r->n * sizeof(double));
    }
    API_END
}

int cf_rci_destroy(cf_rci* r) {
    API_BEGIN
    delete r;
    API_END
}

}  // extern "C"
