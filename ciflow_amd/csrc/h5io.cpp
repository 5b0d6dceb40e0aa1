// HDF5 files of the reference's integrals and sparse matrices, host only
// (built as libciflow_h5.so against the image's HDF5 C library, so the GPU
// library does not depend on it).  The C ABI is include/ciflow_amd_h5.h.
//
//  * Hamiltonian::save / Hamiltonian::read (Hamiltonian.cpp:322-495): group
//    /Data with L, nGroup, orb2irrep, Econst, nup, ndown and overlap (L*L
//    or empty; files written before it existed lack it);
//    TwoIndex::save / read (TwoIndex.cpp:82-202): /TwoIndex/MetaData/
//    IrrepSizes (attributes nGroup, nIrreps) and, per irrep I with orbitals,
//    /TwoIndex/TwoIndexI/"Matrix elements" = the upper triangle of the
//    irrep's block, element (i, j), i <= j, at i + j (j + 1) / 2
//    (TwoIndex.cpp:68-80); FourIndex::save / read: /FourIndex/MetaData/
//    IrrepSizes (attributes nGroup, nIrreps, theTotalSize) and
//    /FourIndex/FourIndexObject/"Matrix elements" = the unique elements in
//    the order in which FourIndex::calcNumberOfUniqueElements assigns their
//    offsets (FourIndex.cpp:57-151, getPtrAllOK :296-315).
//  * UnitaryMatrix::saveU / loadU (UnitaryMatrix.cpp:469-526): /Data/irrep_I
//    = the irrep's square block, column-major, rows the new orbitals
//    (new_a = sum_i U(a, i) old_i, the orientation OrbitalTransform::fillHamCI
//    applies, OrbitalTransform.cpp:64-238).
//  * SparseMatrix_CRS::write_to_file / read_from_file
//    (SparseMatrix_CRS.cpp:484-661): a group with datasets data (f64, with a
//    "size" attribute), col (u64 on disk, "size" attribute), row (u64 on
//    disk) and the scalar n.
//
// The four-index storage is restated from what the allocation loops define:
// for each symmetry block (I_center, I_i, I_k) with I_j = I_center ^ I_i,
// I_l = I_center ^ I_k, I_i <= I_k, I_i <= I_j <= I_l, the canonical
// elements V(i, j, k, l) (irrep-local indices, physics notation
// <ij|kl> = (ik|jl)) follow in the loop order i, k, j, l with the ranges
// below; every canonical element stands for its eight permutations
// (ijkl, jilk, kjil, jkli, ilkj, lijk, klij, lkji -- FourIndex::getPointer).
#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ciflow_amd_h5.h"

namespace {

thread_local std::string g_err;

struct H5Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void chk(long long status, const std::string& what) {
    if (status < 0) throw H5Error("HDF5: " + what);
}

// Number of irreps of the point groups of the reference's Irreps class (C1,
// Ci, C2, Cs, D2, C2v, C2h, D2h); direct products are bitwise XOR.
int n_irreps(int ngroup) {
    static const int n[8] = {1, 2, 2, 2, 4, 4, 4, 8};
    if (ngroup < 0 || ngroup > 7) throw H5Error("nGroup must be in [0, 7]");
    return n[ngroup];
}

// Orbitals of each irrep in increasing global order (the reference's
// orb2indexSy numbering, Hamiltonian.cpp:424-431).
std::vector<std::vector<int>> irrep_orbitals(int L, int nI, const int* orb2irrep) {
    std::vector<std::vector<int>> g(nI);
    for (int p = 0; p < L; ++p) {
        if (orb2irrep[p] < 0 || orb2irrep[p] >= nI) throw H5Error("orbital irrep out of range for the group");
        g[orb2irrep[p]].push_back(p);
    }
    return g;
}

// Visits the canonical four-index elements in storage order:
// f(offset, a, b, c, d) with global orbitals (a b | ...) = V(a, b, c, d).
long long four_index_walk(const std::vector<std::vector<int>>& g,
                          const std::function<void(long long, int, int, int, int)>& f) {
    const int nI = (int)g.size();
    long long off = 0;
    for (int Ic = 0; Ic < nI; ++Ic)
        for (int Ii = 0; Ii < nI; ++Ii) {
            const int Ij = Ic ^ Ii;
            const int si = (int)g[Ii].size(), sj = (int)g[Ij].size();
            if (si == 0 || sj == 0) continue;
            for (int Ik = Ii; Ik < nI; ++Ik) {
                const int Il = Ic ^ Ik;
                const int sk = (int)g[Ik].size(), sl = (int)g[Il].size();
                if (sk == 0 || sl == 0 || !(Ii <= Ij && Ij <= Il)) continue;
                for (int i = 0; i < si; ++i)
                    for (int k = (Ii == Ik ? i : 0); k < sk; ++k)
                        for (int j = (Ic == 0 ? i : 0); j < sj; ++j) {
                            int l0;
                            if (Ic == 0) l0 = (i == j) ? k : (Ii == Ik ? j : 0);
                            else l0 = Ii == Ik ? j : 0;
                            for (int l = l0; l < sl; ++l)
                                f(off++, g[Ii][i], g[Ij][j], g[Ik][k], g[Il][l]);
                        }
            }
        }
    return off;
}

void put8(double* V, int L, int a, int b, int c, int d, double v) {
    auto at = [&](int p, int q, int r, int s) -> double& {
        return V[(((size_t)p * L + q) * L + r) * L + s];
    };
    at(a, b, c, d) = v;
    at(b, a, d, c) = v;
    at(c, b, a, d) = v;
    at(b, c, d, a) = v;
    at(a, d, c, b) = v;
    at(d, a, b, c) = v;
    at(c, d, a, b) = v;
    at(d, c, b, a) = v;
}

// RAII handles
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t), const std::string& what) : id(i), close(c) { chk(i, what); }
    Hid(const Hid&) = delete;
    ~Hid() {
        if (id >= 0 && close) close(id);
    }
    operator hid_t() const { return id; }
};

template <typename T>
void write_1d(hid_t group, const char* name, hid_t ftype, hid_t mtype, const T* data, hsize_t n) {
    Hid space(H5Screate_simple(1, &n, nullptr), H5Sclose, std::string("dataspace ") + name);
    Hid ds(H5Dcreate2(group, name, ftype, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
           std::string("create ") + name);
    if (n) chk(H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), std::string("write ") + name);
}

void attr_scalar(hid_t obj, const char* name, hid_t ftype, hid_t mtype, const void* v) {
    Hid space(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
    Hid a(H5Acreate2(obj, name, ftype, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, std::string("attribute ") + name);
    chk(H5Awrite(a, mtype, v), std::string("write attribute ") + name);
}

hsize_t dataset_len(hid_t ds) {
    Hid space(H5Dget_space(ds), H5Sclose, "dataset space");
    const int rank = H5Sget_simple_extent_ndims(space);
    if (rank == 0) return 1;
    if (rank < 0 || rank > 4) throw H5Error("dataset of unexpected rank");
    hsize_t dims[4] = {0, 0, 0, 0};
    chk(H5Sget_simple_extent_dims(space, dims, nullptr), "dataset dims");
    hsize_t n = 1;
    for (int r = 0; r < rank; ++r) n *= dims[r];
    return n;
}

template <typename T>
std::vector<T> read_all(hid_t group, const char* name, hid_t mtype) {
    Hid ds(H5Dopen2(group, name, H5P_DEFAULT), H5Dclose, std::string("open ") + name);
    std::vector<T> v(dataset_len(ds));
    if (!v.empty()) chk(H5Dread(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), std::string("read ") + name);
    return v;
}

void write_meta(hid_t parent, const std::vector<int>& sizes, int ngroup, const long long* total) {
    Hid meta(H5Gcreate2(parent, "MetaData", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "MetaData");
    const hsize_t n = sizes.size();
    Hid space(H5Screate_simple(1, &n, nullptr), H5Sclose, "IrrepSizes space");
    Hid ds(H5Dcreate2(meta, "IrrepSizes", H5T_STD_I32LE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
           "IrrepSizes");
    chk(H5Dwrite(ds, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, sizes.data()), "write IrrepSizes");
    const int nI = (int)sizes.size();
    attr_scalar(ds, "nGroup", H5T_STD_I32LE, H5T_NATIVE_INT, &ngroup);
    attr_scalar(ds, "nIrreps", H5T_STD_I32LE, H5T_NATIVE_INT, &nI);
    if (total) attr_scalar(ds, "theTotalSize", H5T_STD_I64LE, H5T_NATIVE_LLONG, total);
}

int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}

// Silences the HDF5 library's error stack printing for one call and restores
// the caller's handler afterwards (failures come back as status codes).
struct QuietHdf5 {
    H5E_auto2_t func = nullptr;
    void* data = nullptr;
    QuietHdf5() {
        H5Eget_auto2(H5E_DEFAULT, &func, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietHdf5() { H5Eset_auto2(H5E_DEFAULT, func, data); }
};

}  // namespace

extern "C" {

const char* cf_h5_last_error(void) { return g_err.c_str(); }

int cf_h5_read_hamiltonian(const char* path, int* L_out, int* nup, int* ndown, double* econst, int* ngroup,
                           int* orb2irrep, double* h, double* V, double* overlap, int* has_overlap) {
    try {
        if (!path || !L_out) return fail(1, "NULL argument");
        QuietHdf5 quiet;
        Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, std::string("open ") + path);
        Hid data(H5Gopen2(file, "/Data", H5P_DEFAULT), H5Gclose, "/Data");
        const int L = read_all<int>(data, "L", H5T_NATIVE_INT).at(0);
        const int ng = read_all<int>(data, "nGroup", H5T_NATIVE_INT).at(0);
        *L_out = L;
        if (nup) *nup = read_all<int>(data, "nup", H5T_NATIVE_INT).at(0);
        if (ndown) *ndown = read_all<int>(data, "ndown", H5T_NATIVE_INT).at(0);
        if (econst) *econst = read_all<double>(data, "Econst", H5T_NATIVE_DOUBLE).at(0);
        if (ngroup) *ngroup = ng;
        const std::vector<int> irr = read_all<int>(data, "orb2irrep", H5T_NATIVE_INT);
        if ((int)irr.size() != L) return fail(1, "orb2irrep does not have L entries");
        if (orb2irrep) std::memcpy(orb2irrep, irr.data(), L * sizeof(int));
        // the overlap is optional: absent in older files, empty when the
        // Hamiltonian had none (Hamiltonian.cpp:444-456)
        int ov = 0;
        if (H5Lexists(data, "overlap", H5P_DEFAULT) > 0) {
            const std::vector<double> S = read_all<double>(data, "overlap", H5T_NATIVE_DOUBLE);
            if (S.size() == (size_t)L * L) {
                ov = 1;
                if (overlap) std::memcpy(overlap, S.data(), S.size() * sizeof(double));
            } else if (!S.empty()) {
                return fail(1, "overlap does not have L * L entries");
            }
        }
        if (has_overlap) *has_overlap = ov;
        if (!h && !V) return 0;  // sizes only
        const int nI = n_irreps(ng);
        const auto g = irrep_orbitals(L, nI, irr.data());
        std::vector<int> sizes(nI);
        for (int I = 0; I < nI; ++I) sizes[I] = (int)g[I].size();
        if (h) {
            std::memset(h, 0, (size_t)L * L * sizeof(double));
            Hid two(H5Gopen2(file, "/TwoIndex", H5P_DEFAULT), H5Gclose, "/TwoIndex");
            const std::vector<int> s2 = read_all<int>(Hid(H5Gopen2(two, "MetaData", H5P_DEFAULT), H5Gclose,
                                                          "TwoIndex/MetaData"),
                                                      "IrrepSizes", H5T_NATIVE_INT);
            if (s2 != sizes) return fail(1, "TwoIndex irrep sizes do not match orb2irrep");
            for (int I = 0; I < nI; ++I) {
                const int s = sizes[I];
                if (!s) continue;
                Hid grp(H5Gopen2(two, ("TwoIndex" + std::to_string(I)).c_str(), H5P_DEFAULT), H5Gclose,
                        "TwoIndex" + std::to_string(I));
                const std::vector<double> e = read_all<double>(grp, "Matrix elements", H5T_NATIVE_DOUBLE);
                if ((int)e.size() != s * (s + 1) / 2) return fail(1, "TwoIndex block of the wrong size");
                for (int j = 0; j < s; ++j)
                    for (int i = 0; i <= j; ++i) {
                        const int a = g[I][i], b = g[I][j];
                        h[a * L + b] = h[b * L + a] = e[i + j * (j + 1) / 2];
                    }
            }
        }
        if (V) {
            std::memset(V, 0, (size_t)L * L * L * L * sizeof(double));
            Hid four(H5Gopen2(file, "/FourIndex", H5P_DEFAULT), H5Gclose, "/FourIndex");
            const std::vector<int> s4 = read_all<int>(Hid(H5Gopen2(four, "MetaData", H5P_DEFAULT), H5Gclose,
                                                          "FourIndex/MetaData"),
                                                      "IrrepSizes", H5T_NATIVE_INT);
            if (s4 != sizes) return fail(1, "FourIndex irrep sizes do not match orb2irrep");
            Hid obj(H5Gopen2(four, "FourIndexObject", H5P_DEFAULT), H5Gclose, "FourIndexObject");
            const std::vector<double> e = read_all<double>(obj, "Matrix elements", H5T_NATIVE_DOUBLE);
            const long long n = four_index_walk(g, [](long long, int, int, int, int) {});
            if ((long long)e.size() != n) return fail(1, "FourIndex holds " + std::to_string(e.size()) +
                                                             " elements, the irrep sizes need " + std::to_string(n));
            four_index_walk(g, [&](long long o, int a, int b, int c, int d) { put8(V, L, a, b, c, d, e[o]); });
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(2, e.what());
    }
}

int cf_h5_save_hamiltonian(const char* path, int L, int nup, int ndown, double econst, int ngroup,
                           const int* orb2irrep, const double* h, const double* V, const double* overlap) {
    try {
        if (!path || !h || !V || L < 1) return fail(1, "NULL argument or L < 1");
        QuietHdf5 quiet;
        const int nI = n_irreps(ngroup);
        std::vector<int> irr(L, 0);
        if (orb2irrep) std::memcpy(irr.data(), orb2irrep, L * sizeof(int));
        else if (nI != 1) return fail(1, "orbital irreps are needed for a group other than C1");
        const auto g = irrep_orbitals(L, nI, irr.data());
        std::vector<int> sizes(nI);
        for (int I = 0; I < nI; ++I) sizes[I] = (int)g[I].size();
        // the packed format holds h's upper triangle and one V element per
        // eight permutations: refuse what it would lose
        {
            double hmax = 0.0;
            for (int q = 0; q < L * L; ++q) hmax = std::max(hmax, std::fabs(h[q]));
            const double tol = 1e-12 * std::max(1.0, hmax);  // rotated integrals are symmetric to rounding
            for (int a = 0; a < L; ++a)
                for (int b = 0; b < a; ++b)
                    if (std::fabs(h[a * L + b] - h[b * L + a]) > tol) return fail(1, "h is not symmetric");
        }
        {
            double vmax = 0.0;
            for (size_t q = 0; q < (size_t)L * L * L * L; ++q) vmax = std::max(vmax, std::fabs(V[q]));
            const double tol = 1e-12 * std::max(1.0, vmax);
            auto at = [&](int p, int q, int r, int t) { return V[(((size_t)p * L + q) * L + r) * L + t]; };
            for (int a = 0; a < L; ++a)
                for (int b = 0; b < L; ++b)
                    for (int c = 0; c < L; ++c)
                        for (int d = 0; d < L; ++d) {
                            const double v = at(a, b, c, d);
                            if (std::fabs(at(b, a, d, c) - v) > tol || std::fabs(at(c, b, a, d) - v) > tol ||
                                std::fabs(at(a, d, c, b) - v) > tol)
                                return fail(1, "V lacks the eight-fold permutation symmetry of real orbitals");
                        }
        }
        // the integrals must respect the point group: nothing nonzero outside the stored blocks
        for (int a = 0; a < L; ++a)
            for (int b = 0; b < L; ++b)
                if (irr[a] != irr[b] && std::fabs(h[a * L + b]) > 1e-14)
                    return fail(1, "h couples orbitals of different irreps");
        for (int a = 0; a < L; ++a)
            for (int b = 0; b < L; ++b)
                for (int c = 0; c < L; ++c)
                    for (int d = 0; d < L; ++d)
                        if ((irr[a] ^ irr[b]) != (irr[c] ^ irr[d]) &&
                            std::fabs(V[(((size_t)a * L + b) * L + c) * L + d]) > 1e-14)
                            return fail(1, "V has an element the point group forbids");
        Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, std::string("create ") + path);
        {
            Hid data(H5Gcreate2(file, "/Data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "/Data");
            write_1d(data, "L", H5T_STD_I32LE, H5T_NATIVE_INT, &L, 1);
            write_1d(data, "nGroup", H5T_STD_I32LE, H5T_NATIVE_INT, &ngroup, 1);
            write_1d(data, "orb2irrep", H5T_STD_I32LE, H5T_NATIVE_INT, irr.data(), (hsize_t)L);
            write_1d(data, "Econst", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &econst, 1);
            write_1d(data, "nup", H5T_STD_I32LE, H5T_NATIVE_INT, &nup, 1);
            write_1d(data, "ndown", H5T_STD_I32LE, H5T_NATIVE_INT, &ndown, 1);
            write_1d(data, "overlap", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, overlap, overlap ? (hsize_t)L * L : 0);
        }
        {
            Hid two(H5Gcreate2(file, "TwoIndex", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "TwoIndex");
            write_meta(two, sizes, ngroup, nullptr);
            for (int I = 0; I < nI; ++I) {
                const int s = sizes[I];
                if (!s) continue;
                std::vector<double> e((size_t)s * (s + 1) / 2);
                for (int j = 0; j < s; ++j)
                    for (int i = 0; i <= j; ++i) e[i + j * (j + 1) / 2] = h[g[I][i] * L + g[I][j]];
                Hid grp(H5Gcreate2(two, ("TwoIndex" + std::to_string(I)).c_str(), H5P_DEFAULT, H5P_DEFAULT,
                                   H5P_DEFAULT),
                        H5Gclose, "TwoIndex block");
                write_1d(grp, "Matrix elements", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, e.data(), e.size());
            }
        }
        {
            Hid four(H5Gcreate2(file, "FourIndex", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "FourIndex");
            const long long n = four_index_walk(g, [](long long, int, int, int, int) {});
            write_meta(four, sizes, ngroup, &n);
            std::vector<double> e((size_t)n);
            four_index_walk(g, [&](long long o, int a, int b, int c, int d) {
                e[o] = V[(((size_t)a * L + b) * L + c) * L + d];
            });
            Hid obj(H5Gcreate2(four, "FourIndexObject", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose,
                    "FourIndexObject");
            write_1d(obj, "Matrix elements", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, e.data(), (hsize_t)n);
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(2, e.what());
    }
}

int cf_h5_read_unitary(const char* path, int L, int ngroup, const int* orb2irrep, double* U) {
    try {
        if (!path || !U || L < 1) return fail(1, "NULL argument or L < 1");
        QuietHdf5 quiet;
        const int nI = n_irreps(ngroup);
        std::vector<int> irr(L, 0);
        if (orb2irrep) std::memcpy(irr.data(), orb2irrep, L * sizeof(int));
        else if (nI != 1) return fail(1, "orbital irreps are needed for a group other than C1");
        const auto g = irrep_orbitals(L, nI, irr.data());
        Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, std::string("open ") + path);
        Hid data(H5Gopen2(file, "/Data", H5P_DEFAULT), H5Gclose, "/Data");
        std::memset(U, 0, (size_t)L * L * sizeof(double));
        for (int I = 0; I < nI; ++I) {
            const int s = (int)g[I].size();
            if (!s) continue;
            const std::vector<double> b = read_all<double>(data, ("irrep_" + std::to_string(I)).c_str(),
                                                           H5T_NATIVE_DOUBLE);
            if ((int)b.size() != s * s) return fail(1, "irrep_" + std::to_string(I) + " block of the wrong size");
            for (int a = 0; a < s; ++a)
                for (int i = 0; i < s; ++i) U[(size_t)g[I][a] * L + g[I][i]] = b[a + (size_t)s * i];
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(2, e.what());
    }
}

int cf_h5_save_unitary(const char* path, int L, int ngroup, const int* orb2irrep, const double* U) {
    try {
        if (!path || !U || L < 1) return fail(1, "NULL argument or L < 1");
        QuietHdf5 quiet;
        const int nI = n_irreps(ngroup);
        std::vector<int> irr(L, 0);
        if (orb2irrep) std::memcpy(irr.data(), orb2irrep, L * sizeof(int));
        else if (nI != 1) return fail(1, "orbital irreps are needed for a group other than C1");
        const auto g = irrep_orbitals(L, nI, irr.data());
        for (int a = 0; a < L; ++a)
            for (int i = 0; i < L; ++i)
                if (irr[a] != irr[i] && std::fabs(U[(size_t)a * L + i]) > 1e-14)
                    return fail(1, "U mixes orbitals of different irreps");
        Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, std::string("create ") + path);
        Hid data(H5Gcreate2(file, "/Data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "/Data");
        for (int I = 0; I < nI; ++I) {
            const int s = (int)g[I].size();
            if (!s) continue;
            std::vector<double> b((size_t)s * s);
            for (int a = 0; a < s; ++a)
                for (int i = 0; i < s; ++i) b[a + (size_t)s * i] = U[(size_t)g[I][a] * L + g[I][i]];
            write_1d(data, ("irrep_" + std::to_string(I)).c_str(), H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, b.data(),
                     b.size());
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(2, e.what());
    }
}

int cf_h5_write_crs(const char* path, const char* name, int append, uint64_t n, uint64_t nnz, const uint32_t* row,
                    const uint32_t* col, const double* data) {
    try {
        if (!path || !name || !row || (nnz && (!col || !data))) return fail(1, "NULL argument");
        if (n >= (1ull << 32) || nnz >= (1ull << 32)) return fail(1, "the reference's CRS holds 32-bit sizes");
        if (row[n] != nnz) return fail(1, "row[n] != nnz");
        QuietHdf5 quiet;
        Hid file(append ? H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT)
                        : H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
                 H5Fclose, std::string("open ") + path);
        Hid grp(H5Gcreate2(file, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, std::string("group ") + name);
        const unsigned size = (unsigned)nnz, nn = (unsigned)n;
        {
            const hsize_t d = nnz;
            Hid space(H5Screate_simple(1, &d, nullptr), H5Sclose, "data space");
            Hid ds(H5Dcreate2(grp, "data", H5T_IEEE_F64LE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
                   "data");
            if (nnz) chk(H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "write data");
            attr_scalar(ds, "size", H5T_STD_U64LE, H5T_NATIVE_UINT, &size);
            Hid dc(H5Dcreate2(grp, "col", H5T_STD_U64LE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
                   "col");
            if (nnz) chk(H5Dwrite(dc, H5T_NATIVE_UINT, H5S_ALL, H5S_ALL, H5P_DEFAULT, col), "write col");
            attr_scalar(dc, "size", H5T_STD_U64LE, H5T_NATIVE_UINT, &size);
        }
        write_1d(grp, "row", H5T_STD_U64LE, H5T_NATIVE_UINT, row, (hsize_t)(n + 1));
        {
            Hid sc(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
            Hid ds(H5Dcreate2(grp, "n", H5T_STD_U64LE, sc, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose, "n");
            chk(H5Dwrite(ds, H5T_NATIVE_UINT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &nn), "write n");
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(2, e.what());
    }
}

int cf_h5_read_crs(const char* path, const char* name, uint64_t* n, uint64_t* nnz, uint32_t* row, uint32_t* col,
                   double* data) {
    try {
        if (!path || !name || !n || !nnz) return fail(1, "NULL argument");
        QuietHdf5 quiet;
        Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, std::string("open ") + path);
        Hid grp(H5Gopen2(file, name, H5P_DEFAULT), H5Gclose, std::string("group ") + name);
        unsigned nn = 0, size = 0;
        {
            Hid ds(H5Dopen2(grp, "n", H5P_DEFAULT), H5Dclose, "n");
            chk(H5Dread(ds, H5T_NATIVE_UINT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &nn), "read n");
        }
        {
            Hid ds(H5Dopen2(grp, "data", H5P_DEFAULT), H5Dclose, "data");
            Hid a(H5Aopen(ds, "size", H5P_DEFAULT), H5Aclose, "data size");
            chk(H5Aread(a, H5T_NATIVE_UINT, &size), "read data size");
            if (dataset_len(ds) != (hsize_t)size) return fail(1, "data does not hold its size attribute's count");
            if (data && size) chk(H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "read data");
        }
        *n = nn;
        *nnz = size;
        {
            Hid ds(H5Dopen2(grp, "col", H5P_DEFAULT), H5Dclose, "col");
            if (dataset_len(ds) != (hsize_t)size) return fail(1, "col does not hold data's count");
            if (col && size) chk(H5Dread(ds, H5T_NATIVE_UINT, H5S_ALL, H5S_ALL, H5P_DEFAULT, col), "read col");
        }
        if (row) {
            Hid ds(H5Dopen2(grp, "row", H5P_DEFAULT), H5Dclose, "row");
            if (dataset_len(ds) != (hsize_t)nn + 1) return fail(1, "row does not have n + 1 entries");
            chk(H5Dread(ds, H5T_NATIVE_UINT, H5S_ALL, H5S_ALL, H5P_DEFAULT, row), "read row");
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(2, e.what());
    }
}

}  // extern "C"
