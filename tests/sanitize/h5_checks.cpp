// Checks of the HDF5 file library (ciflow_amd/csrc/h5io.cpp) for the
// AddressSanitizer + UndefinedBehaviorSanitizer build: the reference's
// integral file read (sizes first, then arrays), a save / read round trip,
// the orbital-transformation and CRS files, and the error paths (missing
// file, missing group, refused integrals, bad arguments).
// Usage: h5_checks <tests/golden dir> <scratch dir>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/ciflow_amd_h5.h"

static int failures = 0;

static void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAILED: %s (%s)\n", what, cf_h5_last_error());
        ++failures;
    }
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: h5_checks <golden dir> <scratch dir>\n");
        return 2;
    }
    const std::string gold = argv[1], tmp = argv[2];
    const std::string ao = gold + "/atomicintegralsbeh2_sto3g.h5";
    int L = 0, nup = 0, ndown = 0, ngroup = 0, has_s = -1;
    double econst = 0.0;
    expect(cf_h5_read_hamiltonian(ao.c_str(), &L, &nup, &ndown, &econst, &ngroup, nullptr, nullptr, nullptr, nullptr,
                                  &has_s) == 0,
           "read sizes");
    expect(L == 7 && nup == 3 && ndown == 3 && ngroup == 7 && has_s == 0, "fixture sizes");
    std::vector<int> irr(L);
    std::vector<double> h((size_t)L * L), V((size_t)L * L * L * L);
    expect(cf_h5_read_hamiltonian(ao.c_str(), &L, nullptr, nullptr, nullptr, nullptr, irr.data(), h.data(), V.data(),
                                  nullptr, nullptr) == 0,
           "read arrays");
    // eight-fold symmetry of what was read
    double asym = 0.0;
    for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b)
            for (int c = 0; c < L; ++c)
                for (int d = 0; d < L; ++d) {
                    const double v = V[(((size_t)a * L + b) * L + c) * L + d];
                    asym = std::fmax(asym, std::fabs(v - V[(((size_t)c * L + d) * L + a) * L + b]));
                }
    expect(asym == 0.0, "V symmetric");

    // round trip with an overlap
    const std::string out = tmp + "/h5_roundtrip.h5";
    std::vector<double> S((size_t)L * L, 0.0);
    for (int a = 0; a < L; ++a) S[(size_t)a * L + a] = 1.0;
    expect(cf_h5_save_hamiltonian(out.c_str(), L, nup, ndown, econst, ngroup, irr.data(), h.data(), V.data(),
                                  S.data()) == 0,
           "save");
    std::vector<double> h2(h.size()), V2(V.size()), S2(S.size());
    int L2 = 0, has2 = 0;
    expect(cf_h5_read_hamiltonian(out.c_str(), &L2, nullptr, nullptr, nullptr, nullptr, nullptr, h2.data(),
                                  V2.data(), S2.data(), &has2) == 0,
           "read back");
    expect(L2 == L && has2 == 1 && h2 == h && V2 == V && S2 == S, "round trip exact");

    // orbital transformation file
    const std::string uf = gold + "/unitary-moorthon.h5";
    std::vector<double> U((size_t)L * L);
    expect(cf_h5_read_unitary(uf.c_str(), L, ngroup, irr.data(), U.data()) == 0, "read unitary");
    const std::string uo = tmp + "/h5_unitary.h5";
    expect(cf_h5_save_unitary(uo.c_str(), L, ngroup, irr.data(), U.data()) == 0, "save unitary");
    std::vector<double> U2(U.size());
    expect(cf_h5_read_unitary(uo.c_str(), L, ngroup, irr.data(), U2.data()) == 0 && U2 == U, "unitary round trip");

    // CRS file
    const std::string cf = tmp + "/h5_crs.h5";
    const uint32_t row[4] = {0, 2, 3, 4}, col[4] = {0, 2, 1, 2};
    const double val[4] = {1.5, -0.25, 2.0, 3.0};
    expect(cf_h5_write_crs(cf.c_str(), "m", 0, 3, 4, row, col, val) == 0, "write crs");
    const uint32_t zrow[3] = {0, 0, 0};
    expect(cf_h5_write_crs(cf.c_str(), "empty", 1, 2, 0, zrow, nullptr, nullptr) == 0, "append empty crs");
    uint64_t n = 0, nnz = 0;
    expect(cf_h5_read_crs(cf.c_str(), "m", &n, &nnz, nullptr, nullptr, nullptr) == 0 && n == 3 && nnz == 4,
           "crs sizes");
    uint32_t r2[4], c2[4];
    double v2[4];
    expect(cf_h5_read_crs(cf.c_str(), "m", &n, &nnz, r2, c2, v2) == 0, "read crs");
    bool same = true;
    for (int i = 0; i < 4; ++i) same = same && r2[i] == row[i] && c2[i] == col[i] && v2[i] == val[i];
    expect(same, "crs round trip");
    expect(cf_h5_read_crs(cf.c_str(), "empty", &n, &nnz, r2, nullptr, nullptr) == 0 && n == 2 && nnz == 0,
           "empty crs");

    // error paths: status codes, no crash, no leak
    expect(cf_h5_read_hamiltonian((tmp + "/missing.h5").c_str(), &L2, nullptr, nullptr, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nullptr) == 2,
           "missing file");
    expect(cf_h5_read_crs(cf.c_str(), "absent", &n, &nnz, nullptr, nullptr, nullptr) == 2, "missing group");
    const uint32_t bad_row[4] = {0, 1, 2, 9};
    expect(cf_h5_write_crs(cf.c_str(), "bad", 1, 3, 4, bad_row, col, val) == 1, "row[n] != nnz refused");
    std::vector<double> Vb = V;
    Vb[1] += 0.5;  // breaks the permutation symmetry
    expect(cf_h5_save_hamiltonian(out.c_str(), L, nup, ndown, econst, ngroup, irr.data(), h.data(), Vb.data(),
                                  nullptr) == 1,
           "asymmetric V refused");
    expect(cf_h5_save_hamiltonian(out.c_str(), L, nup, ndown, econst, 9, irr.data(), h.data(), V.data(), nullptr) ==
               2,
           "bad group refused");
    expect(cf_h5_read_hamiltonian(nullptr, &L2, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                  nullptr, nullptr) == 1,
           "NULL path refused");
    if (failures) return 1;
    std::printf("all h5 checks passed\n");
    return 0;
}
