// ciflow_oracle.cpp -- CPU restatement of CIFlow's CI ground-state hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: it restates,
// function by function, the reference algorithms that the GPU library
// replaces, so that tests can compare the GPU results against it.  Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark may
// load it, and only as the checker or the timed CPU baseline.  The GPU library
// never links or calls it.
//
// Each function cites the reference file:line it follows.  Where the reference
// depends on a third-party routine that is absent here (ARPACK dsaupd/dseupd,
// MKL dcsrsymv, LAPACK dgemm/dsyev) the oracle restates the published
// algorithm instead and says so.
//
// Pinning (what checks this file, see tests/):
//   or_first/or_next/or_rank      brute force over all L-bit masks, rank(unrank)=id
//   Slater-Condon + CRS build     brute-force second quantisation on tiny L
//                                 (tests/test_oracle_bruteforce.py), dense eigh
//   or_lanczos                    numpy.linalg.eigh, reference KAT energies
//   RDMs                          brute-force <a+a+aa>, trace invariants,
//                                 energy identity E_dens == E_CI
//   or_gen_integrals              committed SHA-256 of the generated arrays and
//                                 closed-form checks (symmetry, PSD of the ERI)
//   or_read_dets, or_build_crs_dets  the reference's selected-CI KAT (Be cc-pVDZ,
//                                 100 determinants, TestDetFile.cpp:32) and the
//                                 FCI energy identity on the full BeH2 list
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint64_t TYPE;  // Options.h:44-47 (TL: unsigned long)

// ---------------------------------------------------------------------------
// Synthetic integral generator (project specification; the reference has
// none).  splitmix64 stream, Box-Muller normals, fixed draw order:
//   h_ii = -2 + 3 i/(L-1); h_ij (i<j) ~ N(0, 0.05^2), row-major upper;
//   B^P (P = 0..2L-1) symmetric, upper incl. diagonal row-major ~ N(0, 0.1^2),
//   B^P_PP += 0.6 for P < L;  (ij|kl) = sum_P B^P_ij B^P_kl (P ascending);
//   physics V[i][k][j][l] = (ij|kl)  (Hamiltonian.cpp:873 converts the same way).
// ---------------------------------------------------------------------------
namespace {
struct Rng {
    uint64_t s;
    bool have_spare = false;
    double spare = 0.0;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() {  // (0, 1]
        return (double)((next() >> 11) + 1) * (1.0 / 9007199254740992.0);
    }
    double normal() {
        if (have_spare) { have_spare = false; return spare; }
        double u1 = uniform(), u2 = uniform();
        double r = std::sqrt(-2.0 * std::log(u1));
        double a = 6.283185307179586 * u2;
        spare = r * std::sin(a);
        have_spare = true;
        return r * std::cos(a);
    }
};
}  // namespace

extern "C" void or_gen_integrals(int L, uint64_t seed, double* h, double* V) {
    Rng rng(seed);
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) h[i * L + j] = 0.0;
    for (int i = 0; i < L; ++i) h[i * L + i] = -2.0 + 3.0 * (double)i / (double)(L > 1 ? L - 1 : 1);
    for (int i = 0; i < L; ++i)
        for (int j = i + 1; j < L; ++j) {
            double x = 0.05 * rng.normal();
            h[i * L + j] = x;
            h[j * L + i] = x;
        }
    const int np = L * (L + 1) / 2;
    const int nP = 2 * L;
    std::vector<double> B((size_t)nP * L * L, 0.0);
    for (int P = 0; P < nP; ++P) {
        double* b = &B[(size_t)P * L * L];
        for (int i = 0; i < L; ++i)
            for (int j = i; j < L; ++j) {
                double x = 0.1 * rng.normal();
                b[i * L + j] = x;
                b[j * L + i] = x;
            }
        if (P < L) b[P * L + P] += 0.6;
    }
    (void)np;
    const size_t L2 = (size_t)L * L, L3 = L2 * L;
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j)
            for (int k = 0; k < L; ++k)
                for (int l = 0; l < L; ++l) {
                    double s = 0.0;
                    for (int P = 0; P < nP; ++P) {
                        const double* b = &B[(size_t)P * L2];
                        s += b[i * L + j] * b[k * L + l];
                    }
                    // chemist (ij|kl) -> physics <ik|jl>
                    V[(size_t)i * L3 + (size_t)k * L2 + (size_t)j * L + l] = s;
                }
}

// ---------------------------------------------------------------------------
// Permutator_Bit (Permutator.cpp:113-153) and vertex weights
// (CIMethod.cpp:689-747, identical copies in CIDens.cpp:1226-1267).
// ---------------------------------------------------------------------------
extern "C" TYPE or_first_string(int n) {  // Permutator.cpp:146-153
    TYPE b = 1;
    b = b << n;
    b -= 1;
    return b;
}

extern "C" TYPE or_next_string(TYPE v) {  // Permutator.cpp:113-121 (TL branch)
    TYPE t = v | (v - 1);
    return (t + 1) | (((~t & -~t) - 1) >> (__builtin_ctzll(v) + 1));
}

namespace {
struct VW {
    int L, n;
    std::vector<std::vector<int>> vw;
    VW(int L_, int n_) : L(L_), n(n_), vw(L_ + 1, std::vector<int>(n_ + 1, 0)) {
        // setup_vertex_weights, CIMethod.cpp:663-687
        vw[0][0] = 1;
        for (int k = 1; k <= L - n; k++) vw[k][0] = 1;
        for (int m = 1; m <= n; m++)
            for (int k = m; k <= L - (n - m); k++) vw[k][m] = vw[k - 1][m] + vw[k - 1][m - 1];
    }
    unsigned int weight(TYPE s) const {  // determine_weight, CIMethod.cpp:732-747
        unsigned int w = 1;
        int ne = 0;
        for (int k = 0; k < L; k++) {
            TYPE shiftbit = 1;
            if (s & (shiftbit << k)) {
                ne += 1;
                w += vw[k][ne];
            }
        }
        return w - 1;
    }
};

uint64_t binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    uint64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
    return r;
}
}  // namespace

extern "C" uint32_t or_rank(int L, int n, TYPE s) { return VW(L, n).weight(s); }

extern "C" uint64_t or_binom(int n, int k) { return binom(n, k); }

// Enumerate all strings of the space in Gosper order.
extern "C" void or_enumerate(int L, int n, TYPE* out) {
    uint64_t d = binom(L, n);
    TYPE s = or_first_string(n);
    for (uint64_t i = 0; i < d; ++i) {
        out[i] = s;
        s = or_next_string(s);
    }
}

// ---------------------------------------------------------------------------
// Hamiltonian view: dense getTmat/getVmat values (Hamiltonian.cpp:267-320).
// ---------------------------------------------------------------------------
namespace {
struct Ham {
    int L, nup, ndown;
    const double* h;
    const double* V;
    double getTmat(int i, int j) const { return h[i * L + j]; }
    double getVmat(int i, int j, int k, int l) const {
        return V[(((size_t)i * L + j) * L + k) * L + l];
    }
};

// CIMethod.cpp:313-327
void get_orbs(TYPE up, int* array) {
    int i = 0, count = 0;
    while (up != 0) {
        if ((up & 1) == 1) array[i++] = count;
        up = up >> 1;
        count += 1;
    }
}

// CIMethod.cpp:353-380
void arrange_order(int* orbs, int num, TYPE up1, TYPE up2) {
    (void)up2;
    int orb_reorder[4];
    int upin1 = 0, upin2 = 0;
    if (num == 4) { upin1 = 0; upin2 = 2; }
    else if (num == 2) { upin1 = 0; upin2 = 1; }
    TYPE d = 1;
    for (int i = 0; i < num; i++) {
        if (up1 & (d << orbs[i])) orb_reorder[upin1++] = orbs[i];
        else orb_reorder[upin2++] = orbs[i];
    }
    for (int i = 0; i < num; i++) orbs[i] = orb_reorder[i];
}

// CIMethod.cpp:382-396
int get_nstacked(const int* orbs, TYPE up) {
    int sum = 0;
    int small = orbs[0] < orbs[1] ? orbs[0] : orbs[1];
    int big = orbs[0] < orbs[1] ? orbs[1] : orbs[0];
    TYPE shiftbit = 1;
    for (int i = small; i < big; i++)
        if (up & (shiftbit << i)) sum += 1;
    return sum;
}

struct SC {  // Slater-Condon element kernels, CIMethod.cpp:329-487
    const Ham& H;
    explicit SC(const Ham& h) : H(h) {}

    double diag_one(const int* o, int np) const {  // get_diag_cont_one_bitstring :329-341
        double s = 0;
        for (int k = 0; k < np; k++) s += H.getTmat(o[k], o[k]);
        for (int k = 0; k < np - 1; k++)
            for (int l = k + 1; l < np; l++)
                s += H.getVmat(o[k], o[l], o[k], o[l]) - H.getVmat(o[k], o[l], o[l], o[k]);
        return s;
    }
    double jmix(const int* ou, int nu, const int* od, int nd) const {  // :343-351
        double s = 0;
        for (int k = 0; k < nu; k++)
            for (int l = 0; l < nd; l++) s += H.getVmat(ou[k], od[l], ou[k], od[l]);
        return s;
    }
    // diagonal(up1, down1), :398-407.  nu/nd = number of particles in each string.
    double diagonal(TYPE up1, TYPE down1, int nu, int nd) const {
        int ou[64], od[64];
        get_orbs(up1, ou);
        double s = diag_one(ou, nu);
        get_orbs(down1, od);
        s += diag_one(od, nd);
        s += jmix(ou, nu, od, nd);
        return s;
    }
    // one_diff_orbital(up1, up2, down1), :409-427 (n_same = particles of up1's spin)
    double one_diff(TYPE up1, TYPE up2, TYPE down1, int n_same, int n_other) const {
        int orbs[2] = {0, 0}, other[64], down[64];
        get_orbs(up1 ^ up2, orbs);
        double e = H.getTmat(orbs[0], orbs[1]);
        get_orbs(up1 & up2, other);
        for (int z = 0; z < n_same - 1; z++)
            e += H.getVmat(orbs[0], other[z], orbs[1], other[z]) -
                 H.getVmat(orbs[0], other[z], other[z], orbs[1]);
        int diff = get_nstacked(orbs, up1 & up2);
        get_orbs(down1, down);
        for (int z = 0; z < n_other; z++) e += H.getVmat(orbs[0], down[z], orbs[1], down[z]);
        if (diff % 2 != 0) e *= -1;
        return e;
    }
    // two_diff_orbitals(up1, up2), :429-442
    double two_same(TYPE up1, TYPE up2) const {
        int orbs[4];
        get_orbs(up1 ^ up2, orbs);
        arrange_order(orbs, 4, up1, up2);
        int check[2] = {orbs[0], orbs[1]};
        int diff = get_nstacked(check, up1 & up2);
        check[0] = orbs[2];
        check[1] = orbs[3];
        diff += get_nstacked(check, up1 & up2);
        double e = H.getVmat(orbs[0], orbs[1], orbs[2], orbs[3]) -
                   H.getVmat(orbs[0], orbs[1], orbs[3], orbs[2]);
        if (diff % 2 != 0) e *= -1;
        return e;
    }
    // two_diff_orbitals(up1, up2, down1, down2), :444-457
    double two_mixed(TYPE up1, TYPE up2, TYPE down1, TYPE down2) const {
        int orbs[2], dorbs[2];
        get_orbs(up1 ^ up2, orbs);
        get_orbs(down1 ^ down2, dorbs);
        arrange_order(orbs, 2, up1, up2);
        arrange_order(dorbs, 2, down1, down2);
        int diff = get_nstacked(orbs, up1 & up2);
        diff += get_nstacked(dorbs, down1 & down2);
        double e = H.getVmat(orbs[0], dorbs[0], orbs[1], dorbs[1]);
        if (diff % 2 != 0) e *= -1;
        return e;
    }
    // get_ham_element, :459-487
    double element(TYPE up1, TYPE down1, TYPE up2, TYPE down2) const {
        int pu = __builtin_popcountll(up1 ^ up2), pd = __builtin_popcountll(down1 ^ down2);
        if (pd > 4 || pu > 4) return 0;
        if (pu == 4 && pd == 0) return two_same(up1, up2);
        if (pu == 2) {
            if (pd == 0) return one_diff(up1, up2, down1, H.nup, H.ndown);
            if (pd == 2) return two_mixed(up1, up2, down1, down2);
        } else if (pd == 4 && pu == 0) {
            return two_same(down1, down2);
        } else if (pd == 2 && pu == 0) {
            return one_diff(down1, down2, up1, H.ndown, H.nup);
        } else if (pd == 0 && pu == 0) {
            return diagonal(up1, down1, H.nup, H.ndown);
        }
        return 0.;
    }
};

// SparseMatrix_CRS row assembly (SparseMatrix_CRS.cpp:240-322).
struct CRS {
    std::vector<double> data;
    std::vector<unsigned int> col;
    std::vector<unsigned int> row;
    unsigned int n;
    double zero = 1e-14;
    explicit CRS(unsigned int n_) : n(n_) {}
    void NewRow() {  // :314-320
        if (row.size() == (size_t)(n + 1)) return;
        row.push_back((unsigned int)data.size());
    }
    void PushToRow(unsigned int j, double value) {  // :240-272
        if (std::fabs(value) > zero) {
            if (col.empty() || row.back() == col.size() || col.back() < j) {
                data.push_back(value);
                col.push_back(j);
            } else {
                unsigned int begin = row.back();
                for (unsigned int i = begin; i < col.size(); i++) {
                    if (col[i] > j) {
                        col.insert(col.begin() + i, j);
                        data.insert(data.begin() + i, value);
                        break;
                    } else if (col[i] == j) {
                        data[i] += value;
                        if (std::fabs(data[i]) < zero) {
                            data.erase(data.begin() + i);
                            col.erase(col.begin() + i);
                        }
                        break;
                    }
                }
            }
        }
    }
    void PushToRownoadd(unsigned int j, double value) {  // :274-297
        if (std::fabs(value) > zero) {
            if (col.empty() || row.back() == col.size() || col.back() < j) {
                data.push_back(value);
                col.push_back(j);
            } else {
                unsigned int begin = row.back();
                for (unsigned int i = begin; i < col.size(); i++) {
                    if (col[i] > j) {
                        col.insert(col.begin() + i, j);
                        data.insert(data.begin() + i, value);
                        break;
                    } else if (col[i] == j) {
                        data[i] = value;
                        break;
                    }
                }
            }
        }
    }
};

// DOCI::construct_CI_matrix, CIMethod.cpp:611-660
void doci_construct(const Ham& H, double zero, CRS& mat, int start_l, int end_l) {
    SC sc(H);
    const int L = H.L, np = H.nup;
    VW vw(L, np);
    TYPE shiftbit = 1;
    TYPE up1 = or_first_string(np);
    for (int s = 0; s < start_l; ++s) up1 = or_next_string(up1);
    mat.zero = zero;
    for (long i = start_l; i < end_l; i++) {
        mat.NewRow();
        mat.PushToRow((unsigned)i, sc.diagonal(up1, up1, np, np));
        for (int j = 0; j < L; j++)
            if (up1 & (shiftbit << j)) {
                TYPE up_interm = up1 ^ (shiftbit << j);
                for (int k = j + 1; k < L; k++) {
                    if (up1 & (shiftbit << k)) continue;
                    TYPE up2 = up_interm ^ (shiftbit << k);
                    unsigned int index = vw.weight(up2);
                    mat.PushToRow(index, sc.two_mixed(up1, up2, up1, up2));
                }
            }
        up1 = or_next_string(up1);
    }
}

// FCI::construct_CI_matrix, CIMethod.cpp:749-888
void fci_construct(const Ham& H, double zero, CRS& mat, int start_l, int end_l) {
    SC sc(H);
    const int L = H.L, na = H.nup, nb = H.ndown;
    VW vw(L, na), vwd(L, nb);
    const long ddim = (long)binom(L, nb);
    TYPE shiftbit = 1;
    TYPE up1 = or_first_string(na);
    for (int s = 0; s < start_l; ++s) up1 = or_next_string(up1);
    mat.zero = zero;
    for (long i = start_l; i < end_l; i++) {
        TYPE down1 = or_first_string(nb);
        for (long k = 0; k < ddim; k++) {
            mat.NewRow();
            mat.PushToRow((unsigned)(i * ddim + k), sc.diagonal(up1, down1, na, nb));
            for (int j = 0; j < L; j++)
                if (up1 & (shiftbit << j)) {
                    TYPE up_interm = up1 ^ (shiftbit << j);
                    for (int l = j + 1; l < L; l++) {
                        if (up1 & (shiftbit << l)) continue;
                        TYPE up2 = up_interm ^ (shiftbit << l);
                        unsigned int index = vw.weight(up2);
                        mat.PushToRow((unsigned)(index * ddim + k), sc.one_diff(up1, up2, down1, na, nb));
                        for (int m = 0; m < L; m++)
                            if (down1 & (shiftbit << m)) {
                                TYPE down_interm = down1 ^ (shiftbit << m);
                                for (int w = 0; w < L; w++) {
                                    if (down1 & (shiftbit << w)) continue;
                                    TYPE down2 = down_interm ^ (shiftbit << w);
                                    unsigned int indexd = vwd.weight(down2);
                                    mat.PushToRow((unsigned)(index * ddim + indexd),
                                                  sc.two_mixed(up1, up2, down1, down2));
                                }
                            }
                        for (int m = 0; m < L; m++)
                            if ((up2 & (shiftbit << m)) && m != l) {
                                TYPE ups_interm = up2 ^ (shiftbit << m);
                                for (int w = 0; w < L; w++) {
                                    if ((up2 & (shiftbit << w)) || w == j) continue;
                                    TYPE updex = ups_interm ^ (shiftbit << w);
                                    unsigned int indexdex = vw.weight(updex);
                                    if (indexdex > i)
                                        mat.PushToRownoadd((unsigned)(indexdex * ddim + k),
                                                           sc.two_same(up1, updex));
                                }
                            }
                    }
                }
            for (int j = 0; j < L; j++)
                if (down1 & (shiftbit << j)) {
                    TYPE down_interm = down1 ^ (shiftbit << j);
                    for (int l = j + 1; l < L; l++) {
                        if (down1 & (shiftbit << l)) continue;
                        TYPE down2 = down_interm ^ (shiftbit << l);
                        unsigned int index = vwd.weight(down2);
                        mat.PushToRow((unsigned)(i * ddim + index), sc.one_diff(down1, down2, up1, nb, na));
                        for (int m = 0; m < L; m++)
                            if ((down2 & (shiftbit << m)) && m != l) {
                                TYPE downs_interm = down2 ^ (shiftbit << m);
                                for (int w = 0; w < L; w++) {
                                    if ((down2 & (shiftbit << w)) || w == j) continue;
                                    TYPE downdex = downs_interm ^ (shiftbit << w);
                                    unsigned int indexddex = vwd.weight(downdex);
                                    if (indexddex > k)
                                        mat.PushToRownoadd((unsigned)(i * ddim + indexddex),
                                                           sc.two_same(down1, downdex));
                                }
                            }
                    }
                }
            down1 = or_next_string(down1);
        }
        up1 = or_next_string(up1);
    }
}

// CIMethod::build_parallel workload split, CIMethod.cpp:489-535
std::vector<unsigned int> workload_split(unsigned long long nlines, int num_t) {
    unsigned long long num_elems = (nlines * (nlines + 1ull)) / 2;
    unsigned long long size_part = num_elems / num_t + 1;
    std::vector<unsigned int> workload(num_t + 1);
    workload.front() = 0;
    workload.back() = (unsigned int)nlines;
    for (int i = 1; i < num_t; i++) {
        unsigned int num_lines = workload[i - 1];
        unsigned long long ne = 0;
        while (ne < size_part) ne += nlines - num_lines++;
        if (num_lines > nlines) num_lines = (unsigned int)nlines;
        workload[i] = num_lines;
    }
    return workload;
}
}  // namespace

// ---------------------------------------------------------------------------
// CRS build entry point: rows of lines [start_l, end_l) (alpha strings for FCI,
// determinants for DOCI), assembled with num_t threads exactly as
// build_parallel + add_from_list do (CIMethod.cpp:489-561,
// SparseMatrix_CRS.cpp:663-696).  kind: 0 = FCI, 1 = DOCI.
// Returns an opaque handle; query sizes then copy out.
// ---------------------------------------------------------------------------
struct OrCRS {
    std::vector<double> data;
    std::vector<unsigned int> col;
    std::vector<uint64_t> row;  // n+1, absolute within the built slice
    uint64_t nrows = 0;
};

extern "C" void* or_build_crs(int kind, int L, int nup, int ndown, const double* h, const double* V,
                              double zero, int start_l, int end_l, int num_t) {
    Ham H{L, nup, ndown, h, V};
    if (num_t < 1) num_t = 1;
    const long ddim = (long)binom(L, ndown);
    const unsigned long long lines = (unsigned long long)(end_l - start_l);
    std::vector<unsigned int> wl = workload_split(lines, num_t);
    std::vector<std::unique_ptr<CRS>> parts(num_t);
#pragma omp parallel for schedule(static, 1) num_threads(num_t)
    for (int me = 0; me < num_t; ++me) {
        unsigned a = start_l + wl[me], b = start_l + wl[me + 1];
        unsigned rows = (kind == 0) ? (unsigned)((b - a) * ddim) : (b - a);
        parts[me].reset(new CRS(rows));
        if (kind == 0) fci_construct(H, zero, *parts[me], (int)a, (int)b);
        else doci_construct(H, zero, *parts[me], (int)a, (int)b);
    }
    OrCRS* out = new OrCRS();
    size_t total = 0;
    for (auto& p : parts) total += p->data.size();
    out->data.reserve(total);
    out->col.reserve(total);
    for (auto& p : parts) {
        uint64_t prev = out->data.size();
        out->data.insert(out->data.end(), p->data.begin(), p->data.end());
        out->col.insert(out->col.end(), p->col.begin(), p->col.end());
        for (size_t i = 0; i < p->row.size(); ++i) out->row.push_back(prev + p->row[i]);
        p.reset();
    }
    out->row.push_back(out->data.size());
    out->nrows = out->row.size() - 1;
    return out;
}

extern "C" uint64_t or_crs_nrows(void* m) { return ((OrCRS*)m)->nrows; }
extern "C" uint64_t or_crs_nnz(void* m) { return ((OrCRS*)m)->data.size(); }
extern "C" void or_crs_copy(void* m, uint64_t* row, uint32_t* col, double* val) {
    OrCRS* c = (OrCRS*)m;
    if (row) std::memcpy(row, c->row.data(), c->row.size() * sizeof(uint64_t));
    if (col) std::memcpy(col, c->col.data(), c->col.size() * sizeof(uint32_t));
    if (val) std::memcpy(val, c->data.data(), c->data.size() * sizeof(double));
}
extern "C" void or_crs_free(void* m) { delete (OrCRS*)m; }

// ---------------------------------------------------------------------------
// Selected CI from a determinant list.
// ---------------------------------------------------------------------------
// Permutator_File::read_file (Permutator.cpp:179-219): every line without a
// '#' is "UP|DOWN", each a binary string with the highest orbital first
// (strtoull base 2).  Strings of more than 64 orbitals need the reference's
// multi-word TYPE and are rejected here (return -1); lines without '|' are
// skipped (the reference would misparse them).  Returns the number of
// determinants (at most cap are stored).
extern "C" int64_t or_read_dets(const char* path, uint64_t cap, TYPE* up, TYPE* down) {
    FILE* f = std::fopen(path, "r");
    if (!f) return -2;
    char buf[4096];
    int64_t n = 0;
    while (std::fgets(buf, sizeof buf, f)) {
        std::string line(buf);
        while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
        if (line.find('#') != std::string::npos) continue;
        const size_t cut = line.find('|');
        if (cut == std::string::npos) continue;
        if (cut > 64) {
            std::fclose(f);
            return -1;
        }
        const TYPE u = (TYPE)std::strtoull(line.substr(0, cut).c_str(), nullptr, 2);
        const TYPE d = (TYPE)std::strtoull(line.substr(cut + 1).c_str(), nullptr, 2);
        if ((uint64_t)n < cap) {
            up[n] = u;
            down[n] = d;
        }
        ++n;
    }
    std::fclose(f);
    return n;
}

// FCI_File::construct_CI_matrix (CIMethod.cpp:890-912): row i in file order,
// columns j = i..n-1 with get_ham_element(det_i, det_j), pushed in order.
void file_construct(const Ham& H, double zero, CRS& mat, const TYPE* up, const TYPE* down, uint64_t n,
                    uint64_t start_l, uint64_t end_l) {
    SC sc(H);
    mat.zero = zero;
    for (uint64_t i = start_l; i < end_l; i++) {
        mat.NewRow();
        for (uint64_t j = i; j < n; j++) mat.PushToRow((unsigned)j, sc.element(up[i], down[i], up[j], down[j]));
    }
}

// build_parallel + add_from_list for a determinant list (CIMethod.cpp:489-561).
extern "C" void* or_build_crs_dets(int L, int nup, int ndown, const double* h, const double* V, double zero,
                                   uint64_t n, const TYPE* up, const TYPE* down, int num_t) {
    Ham H{L, nup, ndown, h, V};
    if (num_t < 1) num_t = 1;
    std::vector<unsigned int> wl = workload_split(n, num_t);
    std::vector<std::unique_ptr<CRS>> parts(num_t);
#pragma omp parallel for schedule(static, 1) num_threads(num_t)
    for (int me = 0; me < num_t; ++me) {
        const uint64_t a = wl[me], b = wl[me + 1];
        parts[me].reset(new CRS((unsigned)(b - a)));
        file_construct(H, zero, *parts[me], up, down, n, a, b);
    }
    OrCRS* out = new OrCRS();
    for (auto& p : parts) {
        uint64_t prev = out->data.size();
        out->data.insert(out->data.end(), p->data.begin(), p->data.end());
        out->col.insert(out->col.end(), p->col.begin(), p->col.end());
        for (size_t i = 0; i < p->row.size(); ++i) out->row.push_back(prev + p->row[i]);
        p.reset();
    }
    out->row.push_back(out->data.size());
    out->nrows = out->row.size() - 1;
    return out;
}

// Matrix element between two determinants (get_ham_element, CIMethod.cpp:459-487).
extern "C" double or_ham_element(int L, int nup, int ndown, const double* h, const double* V,
                                 TYPE up1, TYPE down1, TYPE up2, TYPE down2) {
    Ham H{L, nup, ndown, h, V};
    return SC(H).element(up1, down1, up2, down2);
}

// ---------------------------------------------------------------------------
// Matrix-vector products.
// ---------------------------------------------------------------------------
// SparseMatrix_CRS_Sym::mvprod, g++ branch (SparseMatrix_CRS.cpp:700-723):
// serial upper-triangle scatter.
extern "C" void or_mvprod_sym(uint64_t n, const uint64_t* row, const uint32_t* col, const double* data,
                              const double* x, double* y) {
    for (uint64_t i = 0; i < n; i++) y[i] = 0;
    for (uint64_t i = 0; i < n; i++)
        for (uint64_t k = row[i]; k < row[i + 1]; k++) {
            y[i] += data[k] * x[col[k]];
            if (i != col[k]) y[col[k]] += data[k] * x[i];
        }
}

// Full-row gather with OpenMP: the open stand-in for the Intel build's
// threaded mkl_cspblas_dcsrsymv (SparseMatrix_CRS.cpp:705-707), same form as
// SparseMatrix_CRS::mvprod (:342-357).
extern "C" void or_mvprod_full(uint64_t n, const uint64_t* row, const uint32_t* col, const double* data,
                               const double* x, double* y, int num_t) {
#pragma omp parallel for schedule(static) num_threads(num_t)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        double s = 0;
        for (uint64_t k = row[i]; k < row[i + 1]; k++) s += data[k] * x[col[k]];
        y[i] = s;
    }
}

// Expand an upper-triangle CRS into full rows (both triangles, columns sorted).
extern "C" void* or_crs_full_from_upper(uint64_t n, const uint64_t* row, const uint32_t* col,
                                        const double* data) {
    OrCRS* f = new OrCRS();
    std::vector<uint64_t> cnt(n + 1, 0);
    for (uint64_t i = 0; i < n; i++)
        for (uint64_t k = row[i]; k < row[i + 1]; k++) {
            cnt[i]++;
            if (col[k] != i) cnt[col[k]]++;
        }
    f->row.assign(n + 1, 0);
    for (uint64_t i = 0; i < n; i++) f->row[i + 1] = f->row[i] + cnt[i];
    f->col.resize(f->row[n]);
    f->data.resize(f->row[n]);
    std::vector<uint64_t> pos(f->row.begin(), f->row.end() - 1);
    // lower part first (columns < i arrive in ascending source-row order)
    for (uint64_t i = 0; i < n; i++)
        for (uint64_t k = row[i]; k < row[i + 1]; k++) {
            uint64_t c = col[k];
            if (c != i) {
                f->col[pos[c]] = (uint32_t)i;
                f->data[pos[c]] = data[k];
                pos[c]++;
            }
        }
    for (uint64_t i = 0; i < n; i++)
        for (uint64_t k = row[i]; k < row[i + 1]; k++) {
            f->col[pos[i]] = col[k];
            f->data[pos[i]] = data[k];
            pos[i]++;
        }
    f->nrows = n;
    return f;
}

// ---------------------------------------------------------------------------
// Lanczos ground state.  ARPACK (dsaupd/dseupd, arpack-ng 3.1.5-3.3.0) is not
// available, so this restates the published thick-restart Lanczos
// (Wu & Simon, SIAM J. Matrix Anal. Appl. 22 (2000) 602) with full
// re-orthogonalisation, using ARPACK's parameters from
// SparseMatrix_CRS.cpp:360-481: which="SA", nev=1, ncv=min(42,n),
// maxiter = 3n restarts, convergence ||r|| <= tol*|theta|.
// The symmetric tridiagonal/arrowhead eigenproblem is solved with cyclic
// Jacobi rotations (Golub & Van Loan, 8.5).
// ---------------------------------------------------------------------------
namespace {
void jacobi_eig(int m, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z) {
    // A: m x m row-major symmetric (destroyed). Z columns = eigenvectors. Sorted ascending.
    Z.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) Z[i * m + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                tot += A[i * m + j] * A[i * m + j];
                if (i != j) off += A[i * m + j] * A[i * m + j];
            }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                double apq = A[p * m + q];
                if (apq == 0.0) continue;
                double app = A[p * m + p], aqq = A[q * m + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; ++k) {
                    double akp = A[k * m + p], akq = A[k * m + q];
                    A[k * m + p] = c * akp - s * akq;
                    A[k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {
                    double apk = A[p * m + k], aqk = A[q * m + k];
                    A[p * m + k] = c * apk - s * aqk;
                    A[q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    double zkp = Z[k * m + p], zkq = Z[k * m + q];
                    Z[k * m + p] = c * zkp - s * zkq;
                    Z[k * m + q] = s * zkp + c * zkq;
                }
            }
    }
    std::vector<int> idx(m);
    for (int i = 0; i < m; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return A[a * m + a] < A[b * m + b]; });
    w.resize(m);
    std::vector<double> Zs((size_t)m * m);
    for (int c = 0; c < m; ++c) {
        w[c] = A[idx[c] * m + idx[c]];
        for (int k = 0; k < m; ++k) Zs[k * m + c] = Z[k * m + idx[c]];
    }
    Z.swap(Zs);
}

double dot(uint64_t n, const double* a, const double* b) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) s += a[i] * b[i];
    return s;
}
}  // namespace

// Operator: full-row CRS (row/col/data) applied with the OpenMP gather.
// Returns 0 on convergence, 1 if maxiter restarts were reached.
extern "C" int or_lanczos(uint64_t n, const uint64_t* row, const uint32_t* col, const double* data,
                          int ncv, double tol, int maxrestart, int num_t, double* eigval, double* eigvec,
                          int* n_matvec, int* n_restart) {
    const int M = (uint64_t)ncv > n ? (int)n : ncv;  // stride of T
    std::vector<double> Vb((size_t)(M + 1) * n);
    std::vector<double> w(n);
    {  // deterministic start vector
        uint64_t s = 12345;
        double nrm = 0;
        for (uint64_t i = 0; i < n; ++i) {
            s = s * 6364136223846793005ULL + 1442695040888963407ULL;
            double x = 1.0 + 0.5 * ((double)(s >> 11) / 9007199254740992.0 - 0.5);
            Vb[i] = x;
            nrm += x * x;
        }
        nrm = std::sqrt(nrm);
        for (uint64_t i = 0; i < n; ++i) Vb[i] /= nrm;
    }
    std::vector<double> T((size_t)M * M, 0.0);
    int k = 0, nmv = 0, it = 0, status = 1, m = M;
    std::vector<double> th, Y;
    double beta_last = 0;
    for (it = 0; it < maxrestart; ++it) {
        m = M;
        for (int j = k; j < M; ++j) {
            double* vj = &Vb[(size_t)j * n];
            or_mvprod_full(n, row, col, data, vj, w.data(), num_t);
            nmv++;
            std::vector<double> c(j + 1);
            for (int pass = 0; pass < 2; ++pass) {
                for (int i = 0; i <= j; ++i) c[i] = dot(n, &Vb[(size_t)i * n], w.data());
                if (pass == 0) T[(size_t)j * M + j] = c[j];
                else T[(size_t)j * M + j] += c[j];
                for (int i = 0; i <= j; ++i) {
                    const double* vi = &Vb[(size_t)i * n];
                    double ci = c[i];
#pragma omp parallel for schedule(static)
                    for (int64_t q = 0; q < (int64_t)n; ++q) w[q] -= ci * vi[q];
                }
            }
            double beta = std::sqrt(dot(n, w.data(), w.data()));
            beta_last = beta;
            double* vn = &Vb[(size_t)(j + 1) * n];
            for (uint64_t q = 0; q < n; ++q) vn[q] = beta > 0 ? w[q] / beta : 0.0;
            if (j + 1 < M) {
                T[(size_t)j * M + j + 1] = beta;
                T[(size_t)(j + 1) * M + j] = beta;
            }
            if (beta == 0 || (uint64_t)(j + 1) == n) {  // invariant subspace
                m = j + 1;
                break;
            }
        }
        std::vector<double> A((size_t)m * m);
        for (int a = 0; a < m; ++a)
            for (int b = 0; b < m; ++b) A[(size_t)a * m + b] = T[(size_t)a * M + b];
        jacobi_eig(m, A, th, Y);
        double res = std::fabs(beta_last * Y[(size_t)(m - 1) * m + 0]);
        bool done = res <= tol * std::fabs(th[0]) || m < M || (uint64_t)m == n;
        int kk = done ? 1 : std::max(1, std::min(m / 2, m - 2));
        std::vector<double> Vn((size_t)kk * n, 0.0);  // V[:,0:kk] = V Y[:,0:kk]
        for (int c2 = 0; c2 < kk; ++c2)
            for (int a = 0; a < m; ++a) {
                double y = Y[(size_t)a * m + c2];
                const double* va = &Vb[(size_t)a * n];
                double* dst = &Vn[(size_t)c2 * n];
#pragma omp parallel for schedule(static)
                for (int64_t q = 0; q < (int64_t)n; ++q) dst[q] += y * va[q];
            }
        *eigval = th[0];
        std::memcpy(eigvec, Vn.data(), n * sizeof(double));
        if (done) {
            status = 0;
            break;
        }
        // thick restart: Ritz vectors + residual direction, arrowhead T
        std::memmove(&Vb[(size_t)kk * n], &Vb[(size_t)m * n], n * sizeof(double));
        std::memcpy(Vb.data(), Vn.data(), (size_t)kk * n * sizeof(double));
        std::fill(T.begin(), T.end(), 0.0);
        for (int c2 = 0; c2 < kk; ++c2) {
            T[(size_t)c2 * M + c2] = th[c2];
            double s = beta_last * Y[(size_t)(m - 1) * m + c2];
            T[(size_t)c2 * M + kk] = s;
            T[(size_t)kk * M + c2] = s;
        }
        k = kk;
    }
    {  // final eigenvalue: Rayleigh quotient of the normalised Ritz vector
        double nrm = std::sqrt(dot(n, eigvec, eigvec));
        for (uint64_t q = 0; q < n; ++q) eigvec[q] /= nrm;
        or_mvprod_full(n, row, col, data, eigvec, w.data(), num_t);
        nmv++;
        *eigval = dot(n, eigvec, w.data());
    }
    {  // sign convention: largest-magnitude component positive
        uint64_t imax = 0;
        for (uint64_t q = 1; q < n; ++q)
            if (std::fabs(eigvec[q]) > std::fabs(eigvec[imax])) imax = q;
        if (eigvec[imax] < 0)
            for (uint64_t q = 0; q < n; ++q) eigvec[q] = -eigvec[q];
    }
    *n_matvec = nmv;
    *n_restart = it;
    return status;
}

// ---------------------------------------------------------------------------
// RDMs: CIDens storage and builders (CIDens.cpp:166-321, 533-932, 1073-1224,
// 1530-1631; Vector2d::index matrix.cpp:674-689).
// Layout of the outputs (exactly the reference's storage):
//   one[2][L(L+1)/2]          packed upper, idx = p*L + q - p(p+1)/2, p <= q
//   two[3][L^2(L^2+1)/2]      three Vector2d(L^2, L^2) packed-symmetric blocks
//                             [0] aa, [1] ab, [2] bb.
// ---------------------------------------------------------------------------
namespace {
struct Dens {
    int L, nup, ndown;
    double* one;  // 2 * L(L+1)/2
    double* two;  // 3 * L2(L2+1)/2 or null
    size_t osz, tsz;
    Dens(int L_, int nu, int nd, double* o, double* t) : L(L_), nup(nu), ndown(nd), one(o), two(t) {
        osz = (size_t)L * (L + 1) / 2;
        size_t L2 = (size_t)L * L;
        tsz = L2 * (L2 + 1) / 2;
    }
    size_t v2index(size_t r, size_t c) const {  // matrix.cpp:674-685
        size_t cols = (size_t)L * L;
        if (r > c) std::swap(r, c);
        return cols * r + c - r * (r + 1) / 2;
    }
    double& t2(int blk, size_t r, size_t c) { return two[blk * tsz + v2index(r, c)]; }
    double t2(int blk, size_t r, size_t c) const { return two[blk * tsz + v2index(r, c)]; }
    // CIDens.cpp:227-252
    void add_two_rdm(int spin, int o1, int o2, int o3, int o4, double value) {
        if (spin == 1) {
            t2(1, (size_t)o1 * L + o2, (size_t)o3 * L + o4) += value;
            return;
        }
        if (o1 == o2 || o3 == o4) return;
        if (o1 > o2) { std::swap(o1, o2); value *= -1.; }
        if (o3 > o4) { std::swap(o3, o4); value *= -1.; }
        t2(spin, (size_t)o1 * L + o2 - (o1 + 1) * (o1 + 2) / 2, (size_t)o3 * L + o4 - (o3 + 1) * (o3 + 2) / 2) += value;
    }
    // CIDens.cpp:166-195
    double get_two_rdm(int spin, int o1, int o2, int o3, int o4) const {
        if (spin == 1) return t2(1, (size_t)o1 * L + o2, (size_t)o3 * L + o4);
        if (o1 == o2 || o3 == o4) return 0.;
        int sign = 1;
        if (o1 > o2) { std::swap(o1, o2); sign *= -1; }
        if (o3 > o4) { std::swap(o3, o4); sign *= -1; }
        return sign * t2(spin, (size_t)o1 * L + o2 - (o1 + 1) * (o1 + 2) / 2, (size_t)o3 * L + o4 - (o3 + 1) * (o3 + 2) / 2);
    }
    // CIDens.cpp:255-262, 276-288
    double get_one_rdm(int spin, int a, int b) const {
        if (a > b) std::swap(a, b);
        return one[spin * osz + (size_t)a * L + b - (size_t)a * (a + 1) / 2];
    }
    void add_one_rdm(int spin, int a, int b, double v) {
        one[spin * osz + (size_t)a * L + b - (size_t)a * (a + 1) / 2] += v;
    }
    // fill_one_dense, CIDens.cpp:676-757
    void fill_one_dense(TYPE b1, TYPE b2, TYPE same, double contr, int up_or_down) {
        int orbs[2], other[64], sameo[64];
        get_orbs(b1 ^ b2, orbs);
        if (up_or_down == 2) { get_orbs(b1 & b2, other); get_orbs(same, sameo); }
        else if (up_or_down == 3) { get_orbs(b1 & b2, other); get_orbs(same, sameo); }
        int diff = get_nstacked(orbs, b1 & b2);
        if (diff % 2 != 0) contr *= -1;
        if (up_or_down < 2) {
            add_one_rdm(up_or_down, orbs[0], orbs[1], contr);
        } else if (up_or_down == 2) {
            for (int i = 0; i < nup - 1; i++) add_two_rdm(0, orbs[0], other[i], orbs[1], other[i], contr);
            for (int i = 0; i < ndown; i++) add_two_rdm(1, orbs[0], sameo[i], orbs[1], sameo[i], contr);
        } else {
            for (int i = 0; i < ndown - 1; i++) add_two_rdm(2, orbs[0], other[i], orbs[1], other[i], contr);
            for (int i = 0; i < nup; i++) add_two_rdm(1, sameo[i], orbs[0], sameo[i], orbs[1], contr);
        }
    }
    // fill_two_dense, :759-778
    void fill_two_dense(TYPE up1, TYPE up2, double contr, int spin) {
        int orbs[4];
        get_orbs(up1 ^ up2, orbs);
        arrange_order(orbs, 4, up1, up2);
        int check[2] = {orbs[0], orbs[1]};
        int diff = get_nstacked(check, up1 & up2);
        check[0] = orbs[2];
        check[1] = orbs[3];
        diff += get_nstacked(check, up1 & up2);
        if (diff % 2 != 0) contr *= -1;
        add_two_rdm(spin, orbs[0], orbs[1], orbs[2], orbs[3], contr);
    }
    // fill_two_dense_upanddown, :780-806
    void fill_two_updown(TYPE up1, TYPE up2, TYPE d1, TYPE d2, double contr) {
        int orbs[2], dorbs[2];
        get_orbs(up1 ^ up2, orbs);
        get_orbs(d1 ^ d2, dorbs);
        arrange_order(orbs, 2, up1, up2);
        arrange_order(dorbs, 2, d1, d2);
        int diff = get_nstacked(orbs, up1 & up2) + get_nstacked(dorbs, d1 & d2);
        if (diff % 2 != 0) contr *= -1;
        add_two_rdm(1, orbs[0], dorbs[0], orbs[1], dorbs[1], contr);
    }
    // fill_diagonal, :808-853
    void fill_diagonal(TYPE up1, TYPE d1, double contr, bool tw) {
        int ou[64], od[64];
        get_orbs(up1, ou);
        get_orbs(d1, od);
        for (int i = 0; i < nup; i++) {
            if (!tw) add_one_rdm(0, ou[i], ou[i], contr);
            else
                for (int j = i + 1; j < nup; j++) add_two_rdm(0, ou[i], ou[j], ou[i], ou[j], contr);
        }
        for (int i = 0; i < ndown; i++) {
            if (!tw) add_one_rdm(1, od[i], od[i], contr);
            else
                for (int j = i + 1; j < ndown; j++) add_two_rdm(2, od[i], od[j], od[i], od[j], contr);
        }
        if (tw)
            for (int i = 0; i < nup; i++)
                for (int j = 0; j < ndown; j++) add_two_rdm(1, ou[i], od[j], ou[i], od[j], contr);
    }
    // fill_density_matrix, :855-881
    void fill_density_matrix(TYPE up1, TYPE d1, TYPE up2, TYPE d2, double contr) {
        int pu = __builtin_popcountll(up1 ^ up2), pd = __builtin_popcountll(d1 ^ d2);
        if (pd > 4 || pu > 4) return;
        if (pu == 4 && pd == 0) fill_two_dense(up1, up2, contr, 0);
        else if (pu == 2) {
            if (pd == 0) fill_one_dense(up1, up2, d1, contr, 2);
            else if (pd == 2) fill_two_updown(up1, up2, d1, d2, contr);
        } else if (pd == 4 && pu == 0) fill_two_dense(d1, d2, contr, 2);
        else if (pd == 2 && pu == 0) fill_one_dense(d1, d2, up1, contr, 3);
        else if (pd == 0 && pu == 0) fill_diagonal(up1, d1, contr, true);
    }
    // fill_density_matrix_one, :883-896
    void fill_density_matrix_one(TYPE up1, TYPE d1, TYPE up2, TYPE d2, double contr) {
        int pu = __builtin_popcountll(up1 ^ up2), pd = __builtin_popcountll(d1 ^ d2);
        if (pu > 2 || pd > 2) return;
        if (pu == 2 && pd == 0) fill_one_dense(up1, up2, d1, contr, 0);
        else if (pu == 0 && pd == 2) fill_one_dense(d1, d2, up1, contr, 1);
        else if (pu == 0 && pd == 0) fill_diagonal(up1, d1, contr, false);
    }
};
}  // namespace

// DensFCI::construct_CI_one_dens / construct_CI_two_dens over all alpha lines
// (CIDens.cpp:1530-1631).  c is the eigenvector in determinant order.
extern "C" void or_rdm_fci(int L, int nup, int ndown, const double* c, double* one, double* two) {
    Dens D(L, nup, ndown, one, two);
    std::memset(one, 0, 2 * D.osz * sizeof(double));
    if (two) std::memset(two, 0, 3 * D.tsz * sizeof(double));
    const unsigned long ud = binom(L, nup), dd = binom(L, ndown);
    // one-dens: ups (incl. equal), :1536-1553
    TYPE up1 = or_first_string(nup);
    for (unsigned long i = 0; i < ud; i++) {
        TYPE up2 = up1;
        for (unsigned long j = i; j < ud; j++) {
            TYPE down1 = or_first_string(ndown);
            for (unsigned long k = 0; k < dd; k++) {
                double contr = c[i * dd + k] * c[dd * j + k];
                D.fill_density_matrix_one(up1, down1, up2, down1, contr);
                down1 = or_next_string(down1);
            }
            up2 = or_next_string(up2);
        }
        up1 = or_next_string(up1);
    }
    // downs with equal ups, no diagonal, :1564-1577
    up1 = or_first_string(nup);
    for (unsigned long i = 0; i < ud; i++) {
        TYPE down1 = or_first_string(ndown);
        for (unsigned long k = 0; k + 1 < dd; k++) {
            TYPE down2 = or_next_string(down1);
            for (unsigned long j = k + 1; j < dd; j++) {
                double contr = c[i * dd + k] * c[dd * i + j];
                D.fill_density_matrix_one(up1, down1, up1, down2, contr);
                down2 = or_next_string(down2);
            }
            down1 = or_next_string(down1);
        }
        up1 = or_next_string(up1);
    }
    if (!two) return;
    // two-dens, :1582-1603 (ups differ)
    up1 = or_first_string(nup);
    for (unsigned long i = 0; i < ud; i++) {
        TYPE up2 = or_next_string(up1);
        for (unsigned long j = i + 1; j < ud; j++) {
            if (__builtin_popcountll(up1 ^ up2) <= 4) {
                TYPE down1 = or_first_string(ndown);
                for (unsigned long k = 0; k < dd; k++) {
                    TYPE down2 = or_first_string(ndown);
                    for (unsigned long l = 0; l < dd; l++) {
                        double contr = c[i * dd + k] * c[dd * j + l];
                        D.fill_density_matrix(up1, down1, up2, down2, contr);
                        down2 = or_next_string(down2);
                    }
                    down1 = or_next_string(down1);
                }
            }
            up2 = or_next_string(up2);
        }
        up1 = or_next_string(up1);
    }
    // ups equal, :1608-1624
    up1 = or_first_string(nup);
    for (unsigned long i = 0; i < ud; i++) {
        TYPE down1 = or_first_string(ndown);
        for (unsigned long k = 0; k < dd; k++) {
            TYPE down2 = down1;
            for (unsigned long l = k; l < dd; l++) {
                double contr = c[i * dd + k] * c[dd * i + l];
                D.fill_density_matrix(up1, down1, up1, down2, contr);
                down2 = or_next_string(down2);
            }
            down1 = or_next_string(down1);
        }
        up1 = or_next_string(up1);
    }
}

// DensDOCI::construct_CI_one_dens / construct_CI_two_dens (CIDens.cpp:1073-1094,1165-1224).
extern "C" void or_rdm_doci(int L, int npair, const double* c, double* one, double* two) {
    Dens D(L, npair, npair, one, two);
    std::memset(one, 0, 2 * D.osz * sizeof(double));
    if (two) std::memset(two, 0, 3 * D.tsz * sizeof(double));
    const unsigned long dim = binom(L, npair);
    VW vw(L, npair);
    TYPE up1 = or_first_string(npair);
    for (unsigned long i = 0; i < dim; i++) {
        for (int j = 0; j < L; j++) {
            double contr = c[i] * c[i];
            TYPE shiftbit = 1;
            if (up1 & (shiftbit << j)) {
                D.add_one_rdm(0, j, j, contr);
                D.add_one_rdm(1, j, j, contr);
            }
        }
        up1 = or_next_string(up1);
    }
    if (!two) return;
    up1 = or_first_string(npair);
    for (unsigned long i = 0; i < dim; i++) {
        double contr = c[i] * c[i];
        for (int j = 0; j < L; j++) {
            TYPE shiftbit = 1;
            if (up1 & (shiftbit << j)) {
                for (int k = j; k < L; k++) {
                    if (up1 & (shiftbit << k)) {
                        if (j != k) {
                            D.add_two_rdm(0, j, k, j, k, contr);
                            D.add_two_rdm(2, j, k, j, k, contr);
                            D.add_two_rdm(1, k, j, k, j, contr);
                        }
                        D.add_two_rdm(1, j, k, j, k, contr);
                    }
                }
                TYPE up_interm = up1 ^ (shiftbit << j);
                for (int k = j + 1; k < L; k++) {
                    if (up1 & (shiftbit << k)) continue;
                    TYPE up2 = up_interm ^ (shiftbit << k);
                    unsigned int index = vw.weight(up2);
                    double contr2 = c[i] * c[index];
                    D.add_two_rdm(1, j, j, k, k, contr2);
                }
            }
        }
        up1 = or_next_string(up1);
    }
}

// Accessors with the reference semantics (get_one_rdm / get_two_rdm).
extern "C" double or_get_one_rdm(int L, const double* one, int spin, int a, int b) {
    Dens D(L, 0, 0, const_cast<double*>(one), nullptr);
    return D.get_one_rdm(spin, a, b);
}
extern "C" double or_get_two_rdm(int L, const double* two, int spin, int a, int b, int c, int d) {
    Dens D(L, 0, 0, nullptr, const_cast<double*>(two));
    return D.get_two_rdm(spin, a, b, c, d);
}

// get_dens_energy = one + two + Econst (CIDens.cpp:898-932)
extern "C" double or_dens_energy(int L, const double* h, const double* V, double econst, const double* one,
                                 const double* two) {
    Dens D(L, 0, 0, const_cast<double*>(one), const_cast<double*>(two));
    double e1 = 0.0;
    for (int i = 0; i < L; i++)
        for (int j = 0; j < L; j++) {
            e1 += h[i * L + j] * D.get_one_rdm(0, i, j);
            e1 += h[i * L + j] * D.get_one_rdm(1, i, j);
        }
    double e2 = 0.0;
    for (int i = 0; i < L; i++)
        for (int j = 0; j < L; j++)
            for (int k = 0; k < L; k++)
                for (int l = 0; l < L; l++) {
                    double v = V[(((size_t)i * L + j) * L + k) * L + l];
                    e2 += v * D.get_two_rdm(0, i, j, k, l);
                    e2 += 2 * v * D.get_two_rdm(1, i, j, k, l);
                    e2 += v * D.get_two_rdm(2, i, j, k, l);
                }
    return e1 + 0.5 * e2 + econst;
}

extern "C" int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Whole ground-state pipeline in one call (for large golden values without
// Python-side copies): build_parallel + add_from_list, expansion to full rows,
// Lanczos.  Returns the Lanczos status; E and the product count via pointers.
extern "C" int or_ground_state(int kind, int L, int nup, int ndown, const double* h, const double* V,
                               double zero, int num_t, int ncv, double tol, double* E, int* n_matvec,
                               uint64_t* nnz_upper, double* eigvec /* may be null */) {
    const int lines = (int)binom(L, nup);
    OrCRS* up = (OrCRS*)or_build_crs(kind, L, nup, ndown, h, V, zero, 0, lines, num_t);
    *nnz_upper = up->data.size();
    OrCRS* full = (OrCRS*)or_crs_full_from_upper(up->nrows, up->row.data(), up->col.data(), up->data.data());
    const uint64_t n = up->nrows;
    delete up;
    std::vector<double> vec(n);
    int nrs = 0;
    int st = or_lanczos(n, full->row.data(), full->col.data(), full->data.data(), ncv, tol, (int)std::min<uint64_t>(3 * n, 1u << 30),
                        num_t, E, vec.data(), n_matvec, &nrs);
    if (eigvec) std::memcpy(eigvec, vec.data(), n * sizeof(double));
    delete full;
    return st;
}

// OrbitalTransform::fillHamCI / rotate_old_to_new (OrbitalTransform.cpp:64-136,
// 182-200) for one C1 block: h' = U h U^T, V' by four quarter transforms in the
// reference's order (first index, last index, third, second), U[a*L+i] = U(a,i).
extern "C" void or_rotate_integrals(int L, const double* U, const double* h, const double* V, double* ho,
                                    double* Vo) {
    const size_t L2 = (size_t)L * L, L3 = L2 * L, L4 = L3 * L;
    std::vector<double> t(L2, 0.0);
    for (int a = 0; a < L; ++a)
        for (int j = 0; j < L; ++j) {
            double s = 0;
            for (int i = 0; i < L; ++i) s += U[a * L + i] * h[i * L + j];
            t[a * L + j] = s;
        }
    for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b) {
            double s = 0;
            for (int j = 0; j < L; ++j) s += t[a * L + j] * U[b * L + j];
            ho[a * L + b] = s;
        }
    std::vector<double> A(V, V + L4), B(L4, 0.0);
    auto idx = [&](int i, int j, int k, int l) { return (((size_t)i * L + j) * L + k) * L + l; };
    // first index: (ijkl) -> (ajkl)
    for (int a = 0; a < L; ++a)
        for (int j = 0; j < L; ++j)
            for (int k = 0; k < L; ++k)
                for (int l = 0; l < L; ++l) {
                    double s = 0;
                    for (int i = 0; i < L; ++i) s += U[a * L + i] * A[idx(i, j, k, l)];
                    B[idx(a, j, k, l)] = s;
                }
    // last index: (ajkl) -> (ajkd)
    for (int a = 0; a < L; ++a)
        for (int j = 0; j < L; ++j)
            for (int k = 0; k < L; ++k)
                for (int d = 0; d < L; ++d) {
                    double s = 0;
                    for (int l = 0; l < L; ++l) s += B[idx(a, j, k, l)] * U[d * L + l];
                    A[idx(a, j, k, d)] = s;
                }
    // third index: (ajkd) -> (ajcd)
    for (int a = 0; a < L; ++a)
        for (int j = 0; j < L; ++j)
            for (int c = 0; c < L; ++c)
                for (int d = 0; d < L; ++d) {
                    double s = 0;
                    for (int k = 0; k < L; ++k) s += A[idx(a, j, k, d)] * U[c * L + k];
                    B[idx(a, j, c, d)] = s;
                }
    // second index: (ajcd) -> (abcd)
    for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b)
            for (int c = 0; c < L; ++c)
                for (int d = 0; d < L; ++d) {
                    double s = 0;
                    for (int j = 0; j < L; ++j) s += B[idx(a, j, c, d)] * U[b * L + j];
                    Vo[idx(a, b, c, d)] = s;
                }
}
