/*
 * ciflow_amd.h -- C ABI of the MI355X (gfx950) CI ground-state library.
 *
 * Every entry point takes plain C types, opaque handles and host or device
 * pointers, and returns an int status (CF_OK on success); cf_last_error()
 * returns the message of the last failure on the calling thread.  No
 * exceptions cross the ABI.  A handle is used from one host thread at a time;
 * different handles may be used concurrently.
 *
 * Each declaration names the reference CIFlow interface it replaces
 * (file:line in the CIFlow sources).
 */
#ifndef CIFLOW_AMD_H
#define CIFLOW_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CF_OK = 0,
    CF_ERR_ARG = 1,
    CF_ERR_HIP = 2,
    CF_ERR_NOMEM = 3,
    CF_ERR_NOCONV = 4,
    CF_ERR_UNSUPPORTED = 5,
    CF_ERR_STATE = 6
};

/* CI space kinds: CIMethod subclasses FCI and DOCI (CIMethod.h:135-162) and
 * FCI_File, selected CI over a determinant list (CIMethod.h:164-174). */
enum { CF_FCI = 0, CF_DOCI = 1, CF_DETS = 2 };

/* Value storage of the CRS matrix and arithmetic of a product. */
enum {
    CF_F64 = 0,     /* FP64 values (the reference's SparseMatrix_CRS data) */
    CF_F32 = 1,     /* FP32-stored values, FP64 accumulation */
    CF_PREC_AUTO = 2 /* served requests (cf_server_submit, cf_solve_batch, cf_request_footprint)
                        only: CF_F32 where the request's operator reads stored matrix values (a
                        stored CRS: the FP32 copy halves the product's bytes), CF_F64 where it
                        does not (the on-the-fly sigma, the block-staged and one-launch DOCI
                        products: measured equal or faster in FP64, the FP32 phase adding the
                        products of the FP64 polish) */
};

typedef struct cf_ctx cf_ctx;
typedef struct cf_ham cf_ham;
typedef struct cf_space cf_space;
typedef struct cf_crs cf_crs;
typedef struct cf_rci cf_rci;

/* ---- library / device context ------------------------------------------ */
int cf_version(void);
const char* cf_last_error(void);
/* One context = one GPU + one stream.  No reference counterpart (the
 * reference is a single-process CPU library). */
int cf_ctx_create(int device, cf_ctx** out);
int cf_ctx_destroy(cf_ctx* ctx);
int cf_ctx_synchronize(cf_ctx* ctx);
/* Per-kernel event timing on the context stream: enable, then read the
 * number of timed launches and their summed duration (ms) for a kernel
 * family: "spmv_f64", "spmv_f32", "crs_count", "crs_fill", "sigma", "rdm". */
int cf_ctx_enable_timing(cf_ctx* ctx, int enable);
int cf_ctx_kernel_time(cf_ctx* ctx, const char* family, int64_t* launches, double* total_ms);
int cf_ctx_reset_timing(cf_ctx* ctx);
/* A plain device-to-device copy of `bytes` (bytes read + bytes written) run
 * `reps` times on the context stream: the reference kernel of known traffic
 * for scaling hardware memory counters.  No reference counterpart. */
int cf_ctx_stream_copy(cf_ctx* ctx, uint64_t bytes, int reps);

/* ---- integrals: replaces Hamiltonian::getTmat/getVmat/getEconst/getL/
 *      getnup/getndown (Hamiltonian.h:80-136, Hamiltonian.cpp:267-320).
 *      h: L*L row-major, h[i*L+j] = getTmat(i,j).
 *      V: L^4 row-major, V[((i*L+j)*L+k)*L+l] = getVmat(i,j,k,l) (physics).
 *      For a model Hamiltonian with _modham set, pass getVmat's values
 *      (zero unless i == j), as Hamiltonian.cpp:311-318 returns them. */
int cf_ham_create(cf_ctx* ctx, int L, int nup, int ndown, const double* h, const double* V, double econst,
                  cf_ham** out);
int cf_ham_destroy(cf_ham* ham);
/* Seeded synthetic integrals (no reference counterpart: the reference reads
 * psi4 dumps).  h_ii = -2+3i/(L-1), h_ij ~ N(0,0.05^2); (ij|kl) = sum_P
 * B^P_ij B^P_kl over 2L symmetric B^P ~ N(0,0.1^2) with B^P_PP += 0.6 (P<L). */
int cf_gen_integrals(int L, uint64_t seed, double* h, double* V);

/* ---- orbital rotation of the integrals: OrbitalTransform::fillHamCI
 *      (OrbitalTransform.cpp:64-238), h' = U h U^T and
 *      V'_abcd = sum U_ai U_bj U_ck U_dl V_ijkl, with U[a*L+i] = U_ai (new
 *      orbital a in old orbital i; the reference's column-major Umx(a,i)),
 *      four cyclic quarter transforms on FP64 MFMA. */
/* Batched, host arrays: h[batch][L*L], V[batch][L^4], U[batch][L*L]. */
int cf_rotate_integrals(cf_ctx* ctx, int L, int batch, const double* h, const double* V, const double* U,
                        double* h_out, double* V_out);
/* No reference counterpart: cf_rotate_integrals for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Device arrays in and out (same layouts),
 * asynchronous on the context stream; the quarter transforms' work array is a
 * 16 MiB per-GPU device array built into the library, so batch*L^4 must not
 * exceed 2,097,152 (CF_ERR_ARG otherwise) and nothing is allocated.  Calls
 * that overlap on one GPU (other contexts or threads) share that array and
 * race: serialise them. */
int cf_rotate_integrals_noalloc(cf_ctx* ctx, int L, int batch, const double* dh, const double* dV,
                                const double* dU, double* dh_out, double* dV_out);
/* Rotated copy of a Hamiltonian kept on the GPU (Econst unchanged). */
int cf_ham_rotate(const cf_ham* src, const double* U, cf_ham** out);
/* Read a Hamiltonian's integrals back (h: L*L, V: L^4; either may be NULL). */
int cf_ham_get(const cf_ham* ham, double* h, double* V);
/* U = exp(K) for antisymmetric K (UnitaryMatrix::updateUnitary,
 * UnitaryMatrix.cpp:229-340); host arrays L*L. */
int cf_expm_skew(int L, const double* K, double* U);
/* Seeded antisymmetric K with K_ij ~ N(0, sigma^2) (i < j), host array L*L.
 * No reference counterpart: stands in for the optimisers' rotation steps. */
int cf_gen_rotation(int L, uint64_t seed, double sigma, double* K);

/* ---- determinant space: Permutator_Bit::get_start_int/permutate_bit
 *      (Permutator.cpp:113-153) and vertex-weight ranking determine_weight
 *      (CIMethod.cpp:663-747).  Determinant order: alpha rank slow, beta
 *      rank fast (CIMethod.cpp:771). */
int cf_space_create(cf_ham* ham, int kind, cf_space** out);
int cf_space_destroy(cf_space* sp);
/* CIMethod::set_ham (CIMethod.cpp:117-126) for a space that stays the same:
 * the space and every matrix built on it take the new Hamiltonian (same
 * context, L, nup, ndown) instead of being rebuilt.  Position-major values
 * (FCI kernel 4) and the DOCI pair table (kernel 5) are refilled in place on
 * the next use -- their layouts do not depend on the integrals -- and the CRS
 * arrays are rebuilt, pattern included, only when something reads them
 * (export, counts, the CRS kernels).  The resident eigenvector is dropped; an
 * on-the-fly sigma of the space must be created again.  The space keeps a
 * pointer to ham: ham must outlive it (or the next set_ham). */
int cf_space_set_ham(cf_space* sp, cf_ham* ham);
int cf_space_dims(const cf_space* sp, uint64_t* n, uint64_t* dim_up, uint64_t* dim_down);
/* ---- integral files: Hamiltonian::read_file (Hamiltonian.cpp:707-885), the
 *      psi4 dump the reference reads ("****  Molecular Integrals" header with
 *      Nalpha, Nbeta, symmetry, nuclear repulsion, L and the orbital irreps,
 *      then "i j value" one-electron and "i j k l value" chemist-notation
 *      two-electron lines).  Host only.  Call with h = V = irreps = NULL to
 *      learn L, then with h[L*L], V[L^4] (physics, V[((i*L+j)*L+k)*L+l] =
 *      getVmat(i,j,k,l), all eight permutations filled) and irreps[L]. */
int cf_read_integrals(const char* path, int* L, int* nup, int* ndown, double* econst, double* h, double* V,
                      int* irreps);
/* ---- selected CI over a determinant list: Permutator_File::read_file
 *      (Permutator.cpp:179-219) and FCI_File (CIMethod.cpp:890-917).
 * Reads the "UP|DOWN" lines of a determinant file (binary strings, highest
 * orbital first; lines holding '#' are comments, lines without '|' are
 * skipped).  Host only.  Stores at most cap determinants, *count = number in
 * the file; strings longer than 64 orbitals are CF_ERR_UNSUPPORTED. */
int cf_read_determinants(const char* path, uint64_t cap, uint64_t* up, uint64_t* down, uint64_t* count);
/* Space of n determinants (alpha string up[i], beta string down[i], in list
 * order; each with ham's nup/ndown electrons in its L orbitals).  cf_crs_build
 * then builds the reference's FCI_File matrix (rows in list order, full rows
 * here); dims report (n, n, 1).  cf_rdm / cf_rdm_dense give DensFILE's RDMs
 * (CIDens.cpp:1468-1526) for L <= 32 and distinct determinants. */
int cf_space_create_dets(cf_ham* ham, uint64_t n, const uint64_t* up, const uint64_t* down, cf_space** out);
/* Unrank strings [first, first+count) of the n-of-L space on the GPU (host out). */
int cf_strings_unrank(cf_ctx* ctx, int L, int nel, uint64_t first, uint64_t count, uint64_t* out);
/* Rank arbitrary n-of-L strings on the GPU (host in/out). */
int cf_strings_rank(cf_ctx* ctx, int L, int nel, const uint64_t* strings, uint64_t count, uint32_t* ranks);

/* ---- Slater-Condon CRS build: FCI/DOCI::construct_CI_matrix via
 *      CIMethod::build_parallel (CIMethod.cpp:489-561, 611-660, 749-888).
 *      The GPU keeps full rows (both triangles); entries with |v| <= zero
 *      are not stored, as SparseMatrix_CRS::PushToRow does (:240-272).
 *      value_prec CF_F32 also stores an FP32 copy of the values.  DOCI
 *      matrices of >= 60,000 rows, whose products run the block-staged kernel
 *      (no CRS needed), build the CRS arrays when something first reads them
 *      (counts, export, a CRS kernel). */
int cf_crs_build(cf_space* sp, int value_prec, double zero, cf_crs** out);
int cf_crs_destroy(cf_crs* m);
/* Dimension and nonzero counts; pass NULL for counts not needed (a deferred
 * CRS is built only when a count is asked for). */
int cf_crs_info(const cf_crs* m, uint64_t* n, uint64_t* nnz_full, uint64_t* nnz_upper);
/* Export rows [row_begin, row_end) of the upper triangle in the reference
 * layout (SparseMatrix_CRS.h:109-114: row pointer, column, value; columns
 * ascending, col >= row).  row has row_end-row_begin+1 entries, relative to
 * the first exported row; col/val must hold the exported count. */
int cf_crs_count_upper(cf_crs* m, uint64_t row_begin, uint64_t row_end, uint64_t* count);
int cf_crs_export_upper(cf_crs* m, uint64_t row_begin, uint64_t row_end, uint64_t* row, uint32_t* col,
                        double* val);
/* The same export with the reference's 32-bit row pointer
 * (std::vector<unsigned int> row, SparseMatrix_CRS.h:111-112): row has
 * row_end-row_begin+1 uint32_t entries; CF_ERR_UNSUPPORTED (nothing written)
 * when the exported count does not fit in 32 bits. */
int cf_crs_export_upper_u32(cf_crs* m, uint64_t row_begin, uint64_t row_end, uint32_t* row, uint32_t* col,
                            double* val);
/* Full-row export (both triangles), same conventions. */
int cf_crs_count_full(cf_crs* m, uint64_t row_begin, uint64_t row_end, uint64_t* count);
int cf_crs_export_full(cf_crs* m, uint64_t row_begin, uint64_t row_end, uint64_t* row, uint32_t* col,
                       double* val);

/* ---- matrix-vector product: SparseMatrix_CRS_Sym::mvprod
 *      (SparseMatrix_CRS.cpp:700-723), the virtual seam that
 *      arpackDiagonalize calls (:446).  y is overwritten. */
int cf_mvprod(cf_crs* m, const double* x, double* y);                       /* host pointers, synchronous */
int cf_mvprod_dev(cf_crs* m, int value_prec, const double* dx, double* dy);  /* device pointers, ctx stream */
int cf_mvprod_prec(cf_crs* m, int value_prec, const double* x, double* y);   /* host pointers, CF_F64 or CF_F32 values */
/* SpMV kernel selection (no reference counterpart): 0 = automatic, 1 = plain
 * row gather, 2 = FCI alpha-block staged (FCI matrices only), 3 = row gather
 * with vectorised column/value streams, 4 = FCI position-major complete rows
 * with LDS-staged x blocks (FCI, d_beta <= 8192; builds an extra value copy;
 * tile_rows > 0 = workgroup size, a multiple of 64 <= 1024, for both value
 * precisions), 5 = DOCI position-major pair-hop columns with the pair
 * elements in LDS (DOCI matrices; 4 bytes per element; tile_rows > 0 =
 * workgroup size), 6 = DOCI block-staged (no stored columns; x blocks of the
 * rows sharing their high orbitals and the pair elements in LDS; tile_rows >
 * 0 = low orbitals w, 1..12, with 1..20 high orbitals), 7 = sliced rows: a
 * copy of the CRS rows in chunks of 64 rows with interleaved entries, so one
 * load reads the same position of 64 neighbouring rows (any matrix; builds a
 * second copy of the columns and values; tile_rows = waves per chunk,
 * 1, 2, 4, 8 or 16, 0 = 16; built with determinant lists in (up, down) order and FCI
 * spaces of >= 8,192 rows with < 400 alpha strings, else on request).  The
 * DOCI copies are built with matrices of >= 40,000 rows, else on request. */
int cf_crs_set_spmv_kernel(cf_crs* m, int kind, int tile_rows);
/* Kernel the next product with value_prec would run (1-7) and its algorithmic
 * HBM bytes (matrix stream + x + y; kernel 6: diagonal + x + y; kernel 7: the
 * stored entries without the chunk padding), the roofline numerator. */
int cf_crs_spmv_bytes(const cf_crs* m, int value_prec, uint64_t* bytes, int* kernel);
/* Launch shape of one SpMV kernel for one value precision (no reference
 * counterpart; the reference's mvprod has no launch shape).  kernel 1: p1 =
 * lanes per row (8, 16, 32, 64); kernel 3: p1 = lanes per row (4, 8, 16,
 * 32), p2 = unroll (1, 2); kernel 4: p1 = workgroup size (multiple of 64,
 * <= 1024, 0 = default), p2 = beta strings per thread (1, 2 or 4, dividing
 * d_beta; 0 = the single-buffered kernel with p1 in 256/512/1024); kernel 5:
 * p1 = workgroup size (multiple of 64, <= 1024, 0 = 256), p2 = lanes per row
 * (1, 2, 4 or 8); kernel 6: p1 = low orbitals w (0 = min(L/2, 12)); kernel 7:
 * p1 = waves per 64-row chunk (1, 2, 4, 8, 16; 0 = 16), p2 = quads in flight
 * per lane (1, 2, or 4 with 4-16 waves; 0 = 2), building the sliced copy if
 * it does not exist.  The shape
 * becomes the automatic choice for value_prec and the matrix returns to
 * automatic selection (kernel 0). */
int cf_crs_set_spmv_shape(cf_crs* m, int value_prec, int kernel, int p1, int p2);
/* One timed SpMV configuration of the autotuner. */
typedef struct {
    int kernel, p1, p2; /* as cf_crs_set_spmv_shape */
    double ms;          /* mean kernel time over the timed launches */
} cf_tune_entry;
/* Autotuner (no reference counterpart): times every launch shape the matrix
 * supports for value_prec (row gather lanes per row; vectorised lanes x
 * unroll; ELL workgroup size x strings per thread when the ELL copy exists;
 * DOCI hop and block shapes; sliced rows, waves x unroll, for determinant
 * lists -- building the sliced copy -- or when that copy exists)
 * with HIP events on the context stream, one warm-up and reps timed launches
 * each, installs the fastest as the automatic choice for value_prec and
 * returns the table (at most cap entries written, *count = number timed). */
int cf_crs_autotune(cf_crs* m, int value_prec, int reps, cf_tune_entry* table, int cap, int* count);

/* ---- ground state: CIMethod::solve -> SparseMatrix_CRS::arpackDiagonalize
 *      (CIMethod.cpp:241-268, SparseMatrix_CRS.cpp:360-481), ARPACK's
 *      parameters: which=SA, nev=1, ncv=min(42,n), maxiter=3n.
 *      Krylov vectors stay in HBM; the host runs the restart logic. */
typedef struct {
    int ncv;            /* Krylov basis size (42 as the reference) */
    int maxrestart;     /* <= 0: 3n as the reference */
    double tol;         /* ||r|| <= tol*|theta| (1e-12) */
    int precision;      /* CF_F64, or CF_F32: FP32-value Lanczos + FP64 polish */
    double tol_f32;     /* convergence of the FP32 phase (1e-8) */
    uint64_t seed;      /* start-vector seed */
    int check_every;    /* > 0: also test convergence every check_every Lanczos steps
                           inside a restart cycle; 0: only when the basis is full, as
                           ARPACK does.  The FP64 polish of a CF_F32 solve always tests
                           every check_every (4 when 0) steps. */
    double tol_energy;  /* > 0: also stop once ||r||^2 / (mu - theta_0) <= tol_energy with
                           mu = theta_1 - ||r_1|| (second Ritz pair): the Kato-Temple
                           estimate of the eigenvalue error (Eh).  It bounds the error
                           once the Krylov space has resolved the second eigenvalue; a
                           partner state closer than that is not seen and the estimate
                           is then optimistic.  0: off */
    /* Called on the solving thread at every convergence test with the products
     * applied so far, the lowest Ritz value and its residual estimate; NULL: none. */
    void (*progress)(void* user, int n_matvec, double theta, double resid);
    void* progress_user;
    /* CF_LANCZOS (default): the thick-restart Lanczos above, ARPACK's method.
     * CF_DAVIDSON (no reference counterpart; for served solves): diagonal-
     * preconditioned Davidson from the Lanczos start vector (8 plain Krylov
     * steps first), basis of min(ncv, 24) vectors restarted to 4 Ritz
     * vectors; same precision phases and stopping tests.  Needs a space with
     * a diagonal (FCI, DOCI).
     * CF_SOLVER_AUTO: CF_DAVIDSON for a space with a diagonal and at least
     * 2,000 determinants, else CF_LANCZOS. */
    int method;
} cf_solve_opts;
enum { CF_LANCZOS = 0, CF_DAVIDSON = 1, CF_SOLVER_AUTO = 2 };

typedef struct {
    double eigval;       /* lowest eigenvalue of the CI matrix (no Econst) */
    double energy;       /* eigval + Econst: CIMethod::get_ci_energy */
    double resid;        /* ||H x - theta x|| of the returned vector (FP64 operator) */
    int n_matvec;        /* FP64 products */
    int n_matvec_f32;    /* FP32-value products */
    int n_restart;
    int status;          /* 0 converged, 1 restart limit reached (ARPACK info=1) */
    double ms;           /* wall time of the solve */
} cf_solve_info;

void cf_solve_opts_default(cf_solve_opts* o);
/* eigvec: host array of n doubles or NULL.  The vector also stays on the GPU
 * as the space's resident eigenvector (NULL: it is not read back at all). */
int cf_solve(cf_crs* m, const cf_solve_opts* opts, double* eigvec, cf_solve_info* info);
/* The resident eigenvector of the last solve on this space (host out, n), with
 * the sign convention of the returned vectors: largest-magnitude component
 * positive. */
int cf_space_eigvec(cf_space* sp, double* eigvec);

/* ---- on-the-fly sigma for FCI spaces too large to store (replaces the
 *      O(dim^2) on-the-fly product of CI_Big::mvprod, CIMethod.cpp:943-984):
 *      sigma = Ha C + C Hb^T (dense same-spin string Hamiltonians, library
 *      GEMMs) + the alpha-beta part from the string single lists.
 *      prec: CF_F64, or CF_F32 to also prepare the FP32 product. */
typedef struct cf_sigma cf_sigma;
int cf_sigma_create(cf_space* sp, int prec, cf_sigma** out);
int cf_sigma_destroy(cf_sigma* s);
/* Workgroup size of the alpha-beta kernel: 128, 256, 512, 768 or 1024 beta strings
 * per workgroup (tuning; no reference counterpart). */
int cf_sigma_set_threads(cf_sigma* s, int threads);
/* Alpha-beta kernel (tuning; no reference counterpart): 1 = stage each alpha
 * action's V slice and C row, 2 = per-Hamiltonian slice tables with the next
 * action's slice and row prefetched into registers (shapes it does not cover
 * run kernel 1), 3 (default) = kernel 2 with packed LDS offsets and exact
 * beta-list lengths (shapes without an instance run kernel 2). */
int cf_sigma_set_kernel(cf_sigma* s, int kernel);
/* ---- CI_Big (CIMethod.h:176-189, CIMethod.cpp:591-609, 943-1116, 1118-1142):
 *      a determinant-list space (cf_space_create_dets, L <= 32) whose matrix
 *      is not stored, or only its leading stored_rows x stored_rows block
 *      (the reference keeps as many leading rows as its memory budget
 *      allows).  The block is built as FCI_File builds its matrix (drop rule
 *      |v| <= 1e-14); every other element is evaluated in each product from
 *      the sorted determinant keys (at most two excitations), without the
 *      drop rule, as CI_Big::mvprod does.  cf_sigma_apply[_dev] and
 *      cf_solve_direct take the handle (FP64 elements in either value
 *      precision; the solver defaults to Lanczos, there is no diagonal);
 *      cf_rdm on the space gives DensFILE's RDMs.  stored_rows = 0: nothing
 *      stored; stored_rows = n: the whole matrix stored. */
int cf_dets_sigma_create(cf_space* sp, uint64_t stored_rows, cf_sigma** out);
/* y = H x, device pointers on the context stream; value_prec CF_F32 uses
 * FP32 operands with FP64 accumulation of the alpha-beta part. */
int cf_sigma_apply_dev(cf_sigma* s, int value_prec, const double* dx, double* dy);
/* y = H x, host pointers, synchronous. */
int cf_sigma_apply(cf_sigma* s, int value_prec, const double* x, double* y);
/* Ground state with the on-the-fly sigma as the operator (same options and
 * result as cf_solve). */
int cf_solve_direct(cf_sigma* s, const cf_solve_opts* opts, double* eigvec, cf_solve_info* info);
/* 1-RDM of a large FCI space from row dot products of the CI vector (no
 * transition-density matrices); c: host vector (n), one[2][L(L+1)/2] in the
 * reference layout (CIDens.cpp:255-288). */
int cf_rdm1_direct(cf_space* sp, const double* c, double* one);

/* ---- reverse-communication Lanczos with the dsaupd_ contract
 *      (matrix.h:42-52; loop at SparseMatrix_CRS.cpp:441-449): on ido == 1
 *      the caller computes workd[ipntr[1]-1 ...] = A * workd[ipntr[0]-1 ...]
 *      (length n) and calls again; ido == 99 ends.  Host vectors. */
int cf_rci_create(uint64_t n, int ncv, double tol, int maxrestart, uint64_t seed, cf_rci** out);
int cf_rci_step(cf_rci* r, int* ido, double* workd, int* ipntr);
/* dseupd_ counterpart: eigenvalue, eigenvector (n), info (0 / 1). */
int cf_rci_result(cf_rci* r, double* eigval, double* eigvec, int* info);
int cf_rci_destroy(cf_rci* r);

/* ---- RDMs: CIMethod::construct_density -> CIDens::construct_density
 *      (CIMethod.cpp:161-171, CIDens.cpp:323-334, DensFCI :1530-1631,
 *      DensDOCI :1073-1224).  Output layout = the reference storage:
 *      one[2][L(L+1)/2]      packed upper (CIDens.cpp:255-288),
 *      two[3][L^2(L^2+1)/2]  Vector2d(L^2,L^2) blocks aa, ab, bb
 *                            (CIDens.cpp:166-252, matrix.cpp:674-689).
 *      two may be NULL (1-RDM only).  c: host CI vector (n). */
int cf_rdm(cf_space* sp, const double* c, double* one, double* two);
/* Dense form: gamma[2][L][L], Gamma[3][L][L][L][L] with
 * Gamma[b][i][j][k][l] = <a+_is a+_jt a_lt a_ks>, b: 0 aa, 1 ab, 2 bb. */
int cf_rdm_dense(cf_space* sp, const double* c, double* gamma, double* Gamma);

/* ---- served stream of rotated-integral solves: the loop the orbital
 *      optimisers run one request at a time (OrbitalOptimization::update_cim
 *      -> CIMethod::set_ham -> build_ci -> solve, SimulatedAnnealing.cpp:83-87,
 *      LocalMinimizer.cpp:73-235, CIMethod.cpp:117-126, 241-268), here as a
 *      queue served continuously by `slots` worker threads with one context
 *      (stream) each on one GPU.  A request is admitted when its estimated HBM
 *      footprint fits the budget still free.  Results are independent of what
 *      else runs (every reduction has a fixed order). */
typedef struct cf_server cf_server;
typedef struct {
    uint64_t tag;       /* returned with the result */
    int kind;           /* CF_FCI / CF_DOCI */
    int L, nup, ndown;
    const double* h;    /* L*L (copied at submit) */
    const double* V;    /* L^4 (copied at submit) */
    const double* U;    /* L*L rotation applied on the GPU first, or NULL */
    double econst;
    int precision;      /* CF_F64, CF_F32 (FP32 operator + FP64 polish) or CF_PREC_AUTO */
    int ncv;            /* <= 0: 42 */
    double tol;         /* <= 0: 1e-12 */
    int check_every;    /* cf_solve_opts.check_every */
    double tol_energy;  /* cf_solve_opts.tol_energy */
    int direct;         /* FCI only: 1 = on-the-fly sigma (cf_solve_direct) instead of a stored CRS */
    int method;         /* cf_solve_opts.method: CF_LANCZOS, CF_DAVIDSON, or CF_SOLVER_AUTO:
                           the server packs DOCI requests of at most 13,312
                           determinants into batched one-launch solves
                           (cf_solve_doci_batch) and takes Davidson for FCI and
                           the larger DOCI ones */
} cf_request;
typedef struct {
    uint64_t tag;
    int status;         /* CF_OK or the failing entry point's status */
    int solve_status;   /* cf_solve_info.status */
    uint64_t n;         /* determinants */
    double energy, eigval, resid;
    int n_matvec, n_matvec_f32;
    double t_submit, t_start, t_end;  /* steady-clock seconds */
    char error[160];
} cf_result;
/* Many small DOCI solves in one launch (the loop above issues them one after
 * another: OrbitalOptimization::update_cim -> set_ham -> build_ci -> solve,
 * SimulatedAnnealing.cpp:83-87): every request whose DOCI space has at most
 * 13,312 determinants (nalpha == nbeta, L <= 32) is solved by one workgroup
 * running FP64 Lanczos on the matrix of DOCI::construct_CI_matrix
 * (CIMethod.cpp:611-660) applied without storing it; the integrals are
 * rotated by U (or taken as they are when U is NULL) in batched FP64-MFMA
 * transforms.  Stops as cf_solve_opts with tol (<= 0: 1e-12), tol_energy and
 * check_every (<= 0: 2); at most 400 products (solve_status 1 beyond).  The
 * precision, ncv, direct and method fields are not used (FP64 throughout).
 * results[i] belongs to reqs[i]; a request that does not qualify gets an
 * error status and leaves the others unaffected.  Synchronous on the context
 * stream; each result is independent of the rest of the batch. */
int cf_solve_doci_batch(cf_ctx* ctx, int count, const cf_request* reqs, cf_result* results);
/* Batched Davidson solves of the same loop (no counterpart in the reference,
 * which solves each request alone: SimulatedAnnealing.cpp:83-87,
 * CIMethod.cpp:241-268): FCI requests on the on-the-fly sigma (direct = 1,
 * the operator of cf_solve_direct) and DOCI requests (the matrix of
 * DOCI::construct_CI_matrix, CIMethod.cpp:611-660, applied block-staged),
 * method CF_SOLVER_AUTO or CF_DAVIDSON, grouped by shape (kind, L, nalpha,
 * nbeta, basis size min(ncv, 24)) and advanced in lock step, up to 32
 * requests per launch.  Each request runs the solver of cf_solve with
 * CF_DAVIDSON: precision CF_F32 an FP32-operand phase (tol_f32 of
 * cf_solve_opts_default, the tol_energy estimate, or a stall) then an FP64
 * polish (CF_PREC_AUTO: FP64 throughout); stops on tol (<= 0: 1e-12) or the tol_energy estimate; integrals
 * rotated by U (NULL: as given).  n_matvec counts the FP64 products,
 * n_matvec_f32 the FP32 ones.  results[i] belongs to reqs[i]; a request the
 * batch does not take gets an error status and leaves the others unaffected.
 * Synchronous on the context stream; each result is the same in any batch
 * and alone. */
int cf_solve_batch(cf_ctx* ctx, int count, const cf_request* reqs, cf_result* results);
/* hbm_budget_bytes <= 0: 80% of the device memory free at creation. */
int cf_server_create(int device, int slots, double hbm_budget_bytes, cf_server** out);
int cf_server_submit(cf_server* s, const cf_request* r);
/* Move up to max finished results (completion order) into out. */
int cf_server_poll(cf_server* s, cf_result* out, int max, int* count);
/* Wait until every submitted request has finished (timeout_s <= 0: no limit). */
int cf_server_drain(cf_server* s, double timeout_s);
int cf_server_stats(cf_server* s, uint64_t* submitted, uint64_t* finished, uint64_t* hbm_in_use,
                    uint64_t* hbm_budget);
/* Stops the workers after the request each is running; unstarted requests are dropped. */
int cf_server_destroy(cf_server* s);
/* The admission estimate of a request (bytes). */
int cf_request_footprint(const cf_request* r, uint64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* CIFLOW_AMD_H */
