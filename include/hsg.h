/* hsg: MI355X-native shift-invert engine, C ABI.
 *
 * Plain C, plain pointers and sizes. Device pointers are HIP device addresses (for example the
 * data_ptr() of a torch tensor on a ROCm device); streams are hipStream_t passed as void*
 * (NULL = the default stream). Every function returns an error code; the message of the last
 * error of the calling thread is hsg_last_error_message().
 *
 * Two layers, each replacing an interface of the reference (hsolve, file:line given per entry):
 *   1. The factorization seam SparseFactorization<T> (include/hsolve/factor.hpp:341-443), which
 *      the reference fills with MUMPS: analyze / factorize / count / solve / inertia.
 *   2. The kernels that the reference's shift-invert operator and Krylov solver run on the CPU:
 *      spmm (include/hsolve/sparse.hpp:151-155), shifted_residual (sparse.hpp:157-173), the
 *      block Gram / update kernels of CGS2 (src/linalg/dense_ops.cpp:752, :883).
 * The Python package hsolve_amd binds these with ctypes and provides hsolve.solve() above them.
 */
#ifndef HSG_H
#define HSG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: 0 = success, then hs::Error::Code in declaration order plus one
 * (include/hsolve/error.hpp:15-25), matching the Python error classes
 * (python/hsolve/__init__.py:48-130). */
enum {
  HSG_OK = 0,
  HSG_INVALID_INPUT = 1,
  HSG_IO = 2,
  HSG_NOT_HERMITIAN = 3,
  HSG_UNSUPPORTED = 4,
  HSG_FACTORIZATION_FAILED = 5,
  HSG_SINGULAR = 6,
  HSG_OUT_OF_MEMORY_BUDGET = 7,
  HSG_NOT_CONVERGED = 8,
  HSG_INTERNAL = 9
};

/* Arithmetic of a numeric factorization (FactorPrecision, include/hsolve/factor.hpp). */
enum { HSG_DOUBLE = 0, HSG_SINGLE = 1 };

/* Library version string, e.g. "0.1.0" (hs::version, python/hsolve/__init__.py:437). */
const char* hsg_version(void);

/* Message of the last failing call on this thread (empty if none). */
const char* hsg_last_error_message(void);

/* ------------------------------------------------------------------------------------------ */
/* Layer 1: sparse LDL^T factorization (replaces SparseFactorization<T>, factor.hpp:341-443).   */

typedef struct hsg_factor hsg_factor;

/* Analysis options (FactorOptions, factor.hpp:248-293; only what this engine uses). */
typedef struct hsg_analysis_options {
  int32_t leaf_size;          /* nested-dissection leaves (default 64) */
  int32_t panel_width;        /* column block of the large-front factorization (default 64) */
  int32_t small_front_max;    /* fronts of order <= this are factorized in LDS (0 = default) */
  int32_t reserved[5];
} hsg_analysis_options;

/* Statistics of an analysis (FactorEstimate, factor.hpp:302-310, plus the supernodal shape). */
typedef struct hsg_analysis_info {
  int64_t n;
  int64_t nnz;               /* pattern entries (both triangles) */
  int64_t supernodes;
  int64_t levels;
  int64_t max_front;
  int64_t max_pivots;        /* largest number of pivot columns of one front */
  double factor_entries;     /* lower-trapezoid entries of L incl. D (INFOG(20)) */
  double flops;              /* sum over fronts of p^3/3 + p^2 c + p c^2 (RINFOG(1) analogue) */
  double panel_bytes_double; /* device bytes of the FP64 factor panels */
  double cb_bytes_double;    /* device bytes of the FP64 contribution-block array: blocks whose
                                lifetimes (own level to parent level) do not overlap share storage */
  double analyze_s;          /* host wall time of ordering + symbolic analysis */
  double ordering_s;
  int64_t small_fronts;      /* fronts handled by the one-workgroup LDS kernel */
  int64_t large_fronts;
  int64_t front_rows;        /* sum over fronts of their order (length of the row lists) */
  double update_flops;       /* flops of the front-update products of one factorization (the
                                matrix-core GEMMs: trailing updates and contribution blocks, lower
                                triangles only); the rest of `flops` is panel factorization work
                                and the fronts factorized whole in LDS */
  double cb_bytes_unshared;  /* what one block per front would take (the sum of all blocks), FP64 */
  double workspace_bytes_double; /* the L D workspace of the large fronts (shared by the levels), FP64 */
} hsg_analysis_info;

/* Inertia of the last factorization (Inertia, factor.hpp:295-300). valid == 0 after a singular
 * factorization (null pivots), as in the reference. */
typedef struct hsg_inertia {
  int64_t neg;
  int64_t zero;
  int64_t pos;
  int32_t valid;
  int32_t two_by_two;        /* 2x2 pivots chosen */
} hsg_inertia;

/* Analyse the pattern of the pencil K(alpha, beta) = alpha A + beta B (replaces
 * SparseFactorization::analyze, factor.hpp:363-371; MUMPS JOB=1 at
 * src/factor/mumps_factorization.cpp:273). The pattern is the union of A's and B's patterns in
 * full symmetric CSR form (host pointers, n rows, rowptr[n] entries). For every pattern entry e,
 * a_index[e] / b_index[e] give the position of that entry in the value arrays of A / B passed to
 * hsg_factorize (or -1 where the matrix has no entry). Symbolic work (nested-dissection ordering,
 * elimination tree, supernodes, front maps) runs on the host; the maps are then uploaded.
 * order (may be NULL): a given elimination order, new -> old (set_pivot_order, factor.hpp:432). */
int hsg_analyze(int64_t n, const int64_t* rowptr, const int32_t* colind, const int64_t* a_index,
                const int64_t* b_index, const int32_t* order, const hsg_analysis_options* opt,
                hsg_factor** out);

int hsg_analysis_get_info(const hsg_factor* f, hsg_analysis_info* info);

/* Elimination order new -> old (n entries) of the analysis (pivot_order, factor.hpp:424-431). */
int hsg_analysis_get_order(const hsg_factor* f, int32_t* order);

/* Front structure of the analysis, for inspection and tests (no counterpart in the reference,
 * whose supernodal structure stays inside MUMPS). Arrays sized by hsg_analysis_info:
 * supernodes entries for first / ncol / m / parent (-1 = root), supernodes + 1 for rowoff, and
 * front_rows for rows (new numbering, ascending per front; the first ncol are the front's own
 * pivot columns). */
int hsg_analysis_get_fronts(const hsg_factor* f, int32_t* first, int32_t* ncol, int32_t* m, int32_t* parent,
                            int64_t* rowoff, int32_t* rows);

/* Numeric multifrontal LDL^T of K = alpha A + beta B on the GPU (replaces factorize,
 * factor.hpp:373-392; MUMPS JOB=2 at mumps_factorization.cpp:326): Bunch-Kaufman 1x1 / 2x2
 * pivots inside each front, inertia read off D. a_values / b_values: device FP64 arrays indexed
 * by a_index / b_index (b_values may be NULL when no b_index is >= 0). precision: HSG_DOUBLE or
 * HSG_SINGLE. null_pivot_tol > 0 enables null-pivot detection (a pivot block with an eigenvalue
 * of magnitude <= null_pivot_tol * max|K_ij| is a null pivot); 0 disables it (then only exactly
 * zero pivots count). *singular receives 1 when null pivots were found (the reference's
 * factorize() == false); the inertia is written to *inertia in either case. Synchronizes the
 * stream before returning. */
int hsg_factorize(hsg_factor* f, double alpha, double beta, const double* a_values, const double* b_values,
                  int32_t precision, double null_pivot_tol, void* stream, hsg_inertia* inertia,
                  int32_t* singular);

/* count (1..HSG_MAX_BATCH) factorizations of the analysed pattern in the same kernel launches,
 * K_z = alphas[z] A_z + betas[z] B_z, with A_z's / B_z's device value arrays at a_values[z] /
 * b_values[z] (host arrays of device pointers; b_values may be NULL, or hold NULL entries, when
 * no b_index is >= 0): the inertia of each into inertias[z], as hsg_factorize computes it for
 * that K_z alone (same pivots, same counts). The factors are kept side by side in the object's
 * numeric storage (count times one factorization's bytes, until hsg_release_factors or the next
 * factorization) and are not solvable: hsg_factor_solve afterwards reports no usable
 * factorization. Synchronizes the stream. It replaces a loop of SparseFactorization::factorize
 * calls for inertia counts over pencils sharing one pattern (the Bloch k-points of a cell, shifts
 * of one pencil), as the reference's helpers count for one analysis (src/driver/helpers.cpp:73). */
#define HSG_MAX_BATCH 64
int hsg_factorize_batch(hsg_factor* f, int32_t count, const double* alphas, const double* betas,
                        const double* const* a_values, const double* const* b_values, int32_t precision,
                        double null_pivot_tol, void* stream, hsg_inertia* inertias);

/* Solve K X = X in place for nrhs columns (column-major, leading dimension ldx >= n) on the
 * device with the last successful factorization (replaces solve, factor.hpp:394-399; MUMPS JOB=3
 * at mumps_factorization.cpp:506). x is FP64; a single-precision factor rounds it to FP32 for
 * the triangular solves (as the reference's single factorization does). Asynchronous on stream. */
int hsg_factor_solve(hsg_factor* f, double* x, int64_t nrhs, int64_t ldx, void* stream);

/* Frees the numerical factors, keeps the analysis (release_factors, factor.hpp:411-418). */
int hsg_release_factors(hsg_factor* f);

void hsg_factor_destroy(hsg_factor* f);

/* Device blocks freed by factorization objects and analyses are kept by the library for reuse
 * (same sizes recur from solve to solve). This returns the idle ones of the current device to the
 * driver and writes the byte count to *bytes (may be NULL). No counterpart in the reference, whose
 * factor memory is MUMPS' own; a serving loop calls it when it goes idle so that other processes
 * on the device can use the memory. The cache holds at most HSG_CACHE_CAP_GB (environment,
 * default 32) GiB of idle blocks per device. */
int hsg_release_cached_memory(int64_t* bytes);

/* A new factorization object over the analysis of src (shared, not recomputed): its own numeric
 * storage, the same pattern, ordering and supernodal plan. The counterpart of the reference's
 * helper processes that each hold the problem's analysis (src/driver/helpers.cpp): here
 * independent factorizations (spectrum slices, certificate counts) run on different streams of
 * one device. src must outlive the clone; a clone is destroyed with hsg_factor_destroy. */
int hsg_factor_clone(const hsg_factor* src, hsg_factor** out);

/* ------------------------------------------------------------------------------------------ */
/* Layer 2: kernels of the shift-invert operator and the block Krylov solver.                  */

/* Y = alpha * M X + beta * Y for CSR M (n x n, device arrays, int64 row pointers, int32 column
 * indices, FP64 values); X, Y: n x p column-major (spmm, sparse.hpp:151-155). */
int hsg_csr_spmm(int64_t n, const int64_t* rowptr, const int32_t* colind, const double* values,
                 const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p, double alpha, double beta,
                 void* stream);

/* Y = alpha * diag(d) X + beta * Y (the diagonal-B case of spmm). */
int hsg_diag_spmm(int64_t n, const double* d, const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p,
                  double alpha, double beta, void* stream);

/* Shifted residual R = B X - (A - sigma B) Y = B (X + sigma Y) - A Y in one sweep of A and B,
 * with per-column FP64 sums |R|^2 and the componentwise backward-error quotient
 * max_i |R_i| / (|B| |X + sigma Y| + |A - sigma B| |Y|)_i  (shifted_residual,
 * sparse.hpp:157-173). B is CSR (b_rowptr != NULL) or diagonal (b_diag != NULL) or identity
 * (both NULL). rnorm2 / omega: device arrays of p doubles (either may be NULL). */
int hsg_shifted_residual(int64_t n, const int64_t* a_rowptr, const int32_t* a_colind, const double* a_values,
                         const int64_t* b_rowptr, const int32_t* b_colind, const double* b_values,
                         const double* b_diag, double sigma, const double* x, int64_t ldx, const double* y,
                         int64_t ldy, double* r, int64_t ldr, int64_t p, double* rnorm2, double* omega,
                         void* stream);

/* C = V^T diag(w) W (k x p, column-major, ldc >= k) for V: n x k, W: n x p (w may be NULL =
 * identity). The Gram / projection kernel of block CGS2 and Rayleigh-Ritz in the B inner product
 * (gram, dense_ops.cpp:883; orthogonalize_block_cgs2, :752). Deterministic: fixed reduction tree. */
int hsg_gram(int64_t n, const double* v, int64_t ldv, int64_t k, const double* w_diag, const double* w,
             int64_t ldw, int64_t p, double* c, int64_t ldc, void* stream);

/* hsg_gram without allocation: the same kernels and the same result bit for bit, with their
 * partial sums in a fixed 16 MiB device array that the library holds on each device it runs on
 * (until the process ends) instead of a workspace it allocates. No counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Returns HSG_INVALID_INPUT when the partial sums of this
 * shape would exceed the array (about p MiB for real V; any k when p <= 16 and k <= 4096). All
 * calls on one device share the array: calls that overlap in time on one device (different
 * streams or threads) race and must be ordered by the caller. */
int hsg_gram_noalloc(int64_t n, const double* v, int64_t ldv, int64_t k, const double* w_diag, const double* w,
                     int64_t ldw, int64_t p, double* c, int64_t ldc, void* stream);

/* W = beta W + alpha V C for V: n x k, C: k x p (device, column-major), W: n x p. The update of
 * block CGS2 and the basis rotation of a thick restart (rotate_in_place,
 * src/eigen/block_lanczos.hpp:96-98). W must not overlap V. */
int hsg_block_update(int64_t n, const double* v, int64_t ldv, int64_t k, const double* c, int64_t ldc, double* w,
                     int64_t ldw, int64_t p, double alpha, double beta, void* stream);

/* out[c] = x_c^T diag(w) y_c for the p columns (w may be NULL = identity); out: device, p doubles.
 * Column norms and Rayleigh quotients of the driver (linalg::dot / norm2 in
 * src/driver/definite.cpp:152-188 and src/ops/shift_invert_operator.cpp:110-113). Deterministic. */
int hsg_column_dots(int64_t n, const double* x, int64_t ldx, const double* y, int64_t ldy, int64_t p,
                    const double* w_diag, double* out, void* stream);

/* One block step of the Krylov basis in one call: block CGS2 of W (n x p) against V (n x k) in the
 * B inner product, W -= V (V^H B W) twice, then two passes of Cholesky-QR within the block,
 * W <- W R_1^-1, W <- W R_2^-1 (the counterpart of BlockOrthonormalizer<T>::run,
 * src/eigen/block_ortho.cpp:164-232, with this engine's order of passes). B is CSR (b_rowptr !=
 * NULL), diagonal (b_diag) or the identity (both NULL); is_complex: interleaved complex blocks.
 * work: device, at least s (p + 2 k p + 2 p^2) + s n p (s n p more for a CSR B) doubles, s = 2 for
 * complex blocks. Host outputs: norms_in (p doubles, the B-norms W had on entry), coef (k x p
 * row-major scalars, the sum of both passes' V^H B W), r (p x p row-major, R = R_2 R_1). status: 0
 * when W is orthonormal; 1 when the block is rank deficient to breakdown_tol or the first
 * Cholesky-QR pass declined (a Gram matrix not numerically positive definite, or diag(R)
 * spreading more than 1e7), W then holding the CGS2 result; 2 when the second pass declined, W
 * holding W R_1^-1. Waits on the stream twice (the host factors the p x p Gram matrices); the last
 * update of W is left queued on the stream. work must stay allocated until the stream has run it. */
int hsg_block_orthonormalize(int32_t is_complex, int64_t n, const double* v, int64_t ldv, int64_t k, double* w,
                             int64_t ldw, int64_t p, const int64_t* b_rowptr, const int32_t* b_colind,
                             const double* b_values, const double* b_diag, double breakdown_tol, double* work,
                             double* norms_in, double* coef, double* r, int32_t* status, void* stream);

/* Residual sums of eigenpair candidates (pair_metrics, src/driver/projection.cpp:53-70): for each of
 * the p columns, out[3c] = ||ax_c - lambda_c bx_c||^2, out[3c+1] = ||bx_c||^2, out[3c+2] = ||x_c||^2,
 * all blocks n x p with leading dimension ld; lambda and out are device arrays. Deterministic. */
int hsg_pair_sums(int64_t n, const double* ax, const double* bx, const double* x, int64_t ld, int64_t p,
                  const double* lambda, double* out, void* stream);

/* Y = a X + b Y for p columns (linalg::axpy, used by the refinement and rescue steps,
 * src/ops/shift_invert_operator.cpp:154-159). */
int hsg_axpby(int64_t n, double a, const double* x, int64_t ldx, double b, double* y, int64_t ldy, int64_t p,
              void* stream);

/* The Lanczos recurrence of the norm estimate (estimate_norm2, src/driver/norms.cpp:60-95) run
 * on the device without reading anything back: M is CSR (rowptr, colind, values) or diagonal
 * (diag, when rowptr is NULL), real or complex (is_complex: interleaved values and vectors).
 * basis: device, steps + 1 vectors of n scalars one after the other, the first one the
 * normalized start vector. For j < steps: w = M v_j, alpha_j = Re v_j^H w, w orthogonalized
 * twice against v_0..v_j (classical Gram-Schmidt), and for j + 1 < steps beta_j = ||w||,
 * v_{j+1} = w / beta_j (0 when beta_j = 0). coef: device, 2 steps doubles, alpha_j at coef[j],
 * beta_j at coef[steps + j]; work: device, 2 (steps + 1) doubles. The caller reads coef once and
 * truncates the tridiagonal at the first breakdown, as the step-by-step loop would stop there. */
int hsg_lanczos_coefficients(int32_t is_complex, int64_t n, const int64_t* rowptr, const int32_t* colind,
                             const double* values, const double* diag, double* basis, int64_t steps, double* coef,
                             double* work, void* stream);

/* Complex Hermitian pencils: the same kernels for T = std::complex<double> (the reference's
 * templates above instantiated for cplx = std::complex<double>, include/hsolve/types.hpp:14).
 * Every matrix value, vector and
 * C array holds interleaved (real, imaginary) doubles, leading dimensions count complex entries,
 * alpha / beta / sigma and the diagonal weights stay real. hsg_zgram forms C = V^H diag(w) W and
 * hsg_zcolumn_dots out[c] = x_c^H diag(w) y_c (p complex values); hsg_zshifted_residual uses
 * complex moduli in the componentwise quotient. hsg_pair_sums and hsg_axpby serve complex blocks
 * unchanged through their real view (2n doubles per column). */
int hsg_zcsr_spmm(int64_t n, const int64_t* rowptr, const int32_t* colind, const double* values,
                  const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p, double alpha, double beta,
                  void* stream);
int hsg_zdiag_spmm(int64_t n, const double* d, const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p,
                   double alpha, double beta, void* stream);
int hsg_zshifted_residual(int64_t n, const int64_t* a_rowptr, const int32_t* a_colind, const double* a_values,
                          const int64_t* b_rowptr, const int32_t* b_colind, const double* b_values,
                          const double* b_diag, double sigma, const double* x, int64_t ldx, const double* y,
                          int64_t ldy, double* r, int64_t ldr, int64_t p, double* rnorm2, double* omega,
                          void* stream);
int hsg_zgram(int64_t n, const double* v, int64_t ldv, int64_t k, const double* w_diag, const double* w,
              int64_t ldw, int64_t p, double* c, int64_t ldc, void* stream);
int hsg_zblock_update(int64_t n, const double* v, int64_t ldv, int64_t k, const double* c, int64_t ldc, double* w,
                      int64_t ldw, int64_t p, double alpha, double beta, void* stream);
int hsg_zcolumn_dots(int64_t n, const double* x, int64_t ldx, const double* y, int64_t ldy, int64_t p,
                     const double* w_diag, double* out, void* stream);

/* ---------------------------------------------------------------- measurement
 * No counterpart in the reference (its timings are host wall-clock phases, SolveReport time_*_s,
 * include/hsolve/solver.hpp). Phase timing of factorizations and solves with HIP events recorded
 * on each call's stream: phase 0 numeric factorization (work = flops of the symbolic analysis),
 * 1 inverse panels built on the first solve after a factorization (flops), 2 forward sweep and
 * 3 backward sweep of a solve (bytes: every stored factor entry once plus the right-hand sides
 * read and written once), 4 the front-update GEMM launches of factorizations, one record per
 * launch (flops of its lower-triangle outputs; recorded only at level 2).
 * hsg_timing_enable(level) clears the records; level 0 stops recording, 1 records phases 0-3,
 * 2 also phase 4;
 * hsg_timing_read synchronizes the recorded events and returns the sums over all calls. */
int hsg_timing_enable(int level);
int hsg_timing_read(int phase, double* ms, int64_t* launches, double* work, int64_t* calls);

#ifdef __cplusplus
}
#endif

#endif /* HSG_H */
