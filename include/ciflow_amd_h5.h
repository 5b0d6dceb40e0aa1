/* ciflow_amd_h5.h -- the reference's HDF5 files, host only.
 *
 * libciflow_h5.so is a separate library linked against the HDF5 C library,
 * so the GPU library does not need HDF5.  Functions return 0 on success,
 * 1 for bad arguments or a malformed file, 2 for an HDF5 failure; the
 * message is in cf_h5_last_error() (per thread).
 *
 * Integrals: dense row-major h[L*L] and physics-notation
 * V[L*L*L*L] (V[((i*L+j)*L+k)*L+l] = <ij|kl> = getVmat(i,j,k,l)), the same
 * arrays cf_ham_create takes (include/ciflow_amd.h).
 */
#ifndef CIFLOW_AMD_H5_H
#define CIFLOW_AMD_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* cf_h5_last_error(void);

/* Hamiltonian::read (Hamiltonian.cpp:402-495, TwoIndex.cpp:147-202,
 * FourIndex.cpp:379-433).  With h == V == NULL only the sizes and scalars
 * are returned (call first to size the buffers).  orb2irrep receives L
 * entries, overlap L*L when the file holds one (*has_overlap = 1); any
 * output pointer may be NULL.  The symmetry-packed TwoIndex / FourIndex
 * blocks are expanded to the full dense arrays (every element of the eight
 * permutations of each stored one; elements the point group forbids are 0). */
int cf_h5_read_hamiltonian(const char* path, int* L, int* nup, int* ndown, double* econst, int* ngroup,
                           int* orb2irrep, double* h, double* V, double* overlap, int* has_overlap);

/* Hamiltonian::save (Hamiltonian.cpp:322-400, TwoIndex.cpp:82-145,
 * FourIndex.cpp:317-377) to exactly `path` (the reference prefixes "ham"
 * and its short file name and appends ".h5"; callers choose the name here).
 * ngroup 0-7 numbers C1, Ci, C2, Cs, D2, C2v, C2h, D2h; orb2irrep may be
 * NULL for C1; overlap (L*L) may be NULL, then an empty overlap dataset is
 * written as the reference does for a Hamiltonian without one.  Integrals
 * that couple orbitals the point group keeps apart are refused (1): the
 * packed format has no place for them. */
int cf_h5_save_hamiltonian(const char* path, int L, int nup, int ndown, double econst, int ngroup,
                           const int* orb2irrep, const double* h, const double* V, const double* overlap);

/* UnitaryMatrix::loadU / saveU (UnitaryMatrix.cpp:469-526): the per-irrep
 * blocks of an orbital transformation as one dense row-major U[L*L] with
 * U[a*L+i] the coefficient of old orbital i in new orbital a -- the U that
 * cf_ham_rotate applies (OrbitalTransform::fillHamCI).  orb2irrep (L
 * entries) places the blocks; it may be NULL for C1.  save refuses a U that
 * mixes irreps. */
int cf_h5_read_unitary(const char* path, int L, int ngroup, const int* orb2irrep, double* U);
int cf_h5_save_unitary(const char* path, int L, int ngroup, const int* orb2irrep, const double* U);

/* SparseMatrix_CRS::write_to_file (SparseMatrix_CRS.cpp:484-571): group
 * `name` in `path` (append != 0 opens an existing file) with the upper CRS
 * of cf_crs_export_upper_u32 -- row[n+1], col[nnz], data[nnz].  Sizes must
 * fit the reference's 32-bit counters. */
int cf_h5_write_crs(const char* path, const char* name, int append, uint64_t n, uint64_t nnz, const uint32_t* row,
                    const uint32_t* col, const double* data);

/* SparseMatrix_CRS::read_from_file (SparseMatrix_CRS.cpp:578-661): with
 * row == col == data == NULL returns n and nnz only. */
int cf_h5_read_crs(const char* path, const char* name, uint64_t* n, uint64_t* nnz, uint32_t* row, uint32_t* col,
                   double* data);

#ifdef __cplusplus
}
#endif
#endif
