// Seeded synthetic integral generator (the reference has none; it reads psi4
// dumps, Hamiltonian.cpp:707-885).  Specification:
//   splitmix64 stream seeded with `seed`; uniforms u = ((x >> 11) + 1) / 2^53
//   in (0,1]; Box-Muller normals, cos branch first then sin branch.
//   h_ii = -2 + 3 i/(L-1); h_ij (i<j) ~ N(0, 0.05^2) drawn row-major, mirrored.
//   B^P, P = 0..2L-1: symmetric, upper triangle incl. diagonal drawn
//   row-major ~ N(0, 0.1^2), mirrored; B^P_PP += 0.6 for P < L.
//   (ij|kl) = sum_{P ascending} B^P_ij B^P_kl, physics V[i][k][j][l] = (ij|kl)
//   (the chemist -> physics mapping of Hamiltonian.cpp:873).
// The ERI tensor is positive semidefinite and exactly 8-fold symmetric in
// floating point (the products commute and the sum order is fixed).
#include <cmath>
#include <cstdint>
#include <vector>

#pragma clang fp contract(off)

namespace cf {
namespace {
struct SplitMix {
    uint64_t s;
    bool spare_ok = false;
    double spare = 0.0;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }
    double normal() {
        if (spare_ok) {
            spare_ok = false;
            return spare;
        }
        const double u1 = uniform(), u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double a = 6.283185307179586 * u2;
        spare = r * std::sin(a);
        spare_ok = true;
        return r * std::cos(a);
    }
};
}  // namespace

void gen_integrals(int L, uint64_t seed, double* h, double* V) {
    SplitMix g(seed);
    for (int i = 0; i < L * L; ++i) h[i] = 0.0;
    for (int i = 0; i < L; ++i) h[i * L + i] = -2.0 + 3.0 * (double)i / (double)(L > 1 ? L - 1 : 1);
    for (int i = 0; i < L; ++i)
        for (int j = i + 1; j < L; ++j) {
            const double x = 0.05 * g.normal();
            h[i * L + j] = x;
            h[j * L + i] = x;
        }
    const int nP = 2 * L;
    const size_t L2 = (size_t)L * L, L3 = L2 * L;
    std::vector<double> B((size_t)nP * L2, 0.0);
    for (int P = 0; P < nP; ++P) {
        double* b = &B[(size_t)P * L2];
        for (int i = 0; i < L; ++i)
            for (int j = i; j < L; ++j) {
                const double x = 0.1 * g.normal();
                b[i * L + j] = x;
                b[j * L + i] = x;
            }
        if (P < L) b[P * L + P] += 0.6;
    }
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j)
            for (int k = 0; k < L; ++k)
                for (int l = 0; l < L; ++l) {
                    double s = 0.0;
                    for (int P = 0; P < nP; ++P) {
                        const double* b = &B[(size_t)P * L2];
                        const double prod = b[i * L + j] * b[k * L + l];
                        s += prod;
                    }
                    V[(size_t)i * L3 + (size_t)k * L2 + (size_t)j * L + l] = s;
                }
}

}  // namespace cf

namespace cf {

// Antisymmetric generator of an orbital rotation: K_ij ~ N(0, sigma^2) for
// i < j drawn row-major from the splitmix64/Box-Muller stream seeded with
// seed ^ 0x6A09E667F3BCC909, K_ji = -K_ij, K_ii = 0.
void gen_rotation(int L, uint64_t seed, double sigma, double* K) {
    SplitMix g(seed ^ 0x6A09E667F3BCC909ULL);
    for (int i = 0; i < L * L; ++i) K[i] = 0.0;
    for (int i = 0; i < L; ++i)
        for (int j = i + 1; j < L; ++j) {
            const double x = sigma * g.normal();
            K[i * L + j] = x;
            K[j * L + i] = -x;
        }
}

// U = exp(K) by scaling and squaring with a degree-18 Taylor polynomial
// (Moler & Van Loan, "Nineteen dubious ways", method 3): ||K/2^s||_1 <= 1/4,
// truncation error < (1/4)^19/19! ~ 1e-28.  The reference builds exp(X) from
// the eigenvectors of X^2 and 2x2 rotation blocks (UnitaryMatrix.cpp:229-340);
// both give the same orthogonal matrix to rounding.
void expm_skew(int L, const double* K, double* U) {
    double nrm = 0.0;
    for (int j = 0; j < L; ++j) {
        double c = 0.0;
        for (int i = 0; i < L; ++i) c += std::fabs(K[i * L + j]);
        nrm = nrm > c ? nrm : c;
    }
    int s = 0;
    while (nrm > 0.25) {
        nrm *= 0.5;
        ++s;
    }
    const double scale = std::ldexp(1.0, -s);
    std::vector<double> A((size_t)L * L), T((size_t)L * L), R((size_t)L * L), tmp((size_t)L * L);
    for (int i = 0; i < L * L; ++i) A[i] = K[i] * scale;
    for (int i = 0; i < L * L; ++i) R[i] = T[i] = (i % (L + 1) == 0) ? 1.0 : 0.0;
    auto mul = [L](const std::vector<double>& X, const std::vector<double>& Y, std::vector<double>& Z) {
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < L; ++j) {
                double acc = 0.0;
                for (int k = 0; k < L; ++k) acc += X[i * L + k] * Y[k * L + j];
                Z[i * L + j] = acc;
            }
    };
    for (int k = 1; k <= 18; ++k) {
        mul(T, A, tmp);
        for (int i = 0; i < L * L; ++i) T[i] = tmp[i] / (double)k;
        for (int i = 0; i < L * L; ++i) R[i] += T[i];
    }
    for (int q = 0; q < s; ++q) {
        mul(R, R, tmp);
        R.swap(tmp);
    }
    for (int i = 0; i < L * L; ++i) U[i] = R[i];
}

}  // namespace cf
