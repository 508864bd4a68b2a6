/* hsg_solve: the whole eigensolve behind one C call.
 *
 * Replaces the reference's upper boundary, the extension entry
 *   _hsolve.solve(a, b, config_json, verbose, storage)
 *     -> (w, v, complex_w, complex_v, report_json, config_json, complex_dtype)
 * (python/src/module.cpp:326-345, bound at :458-461), which python/hsolve/__init__.py calls and
 * which runs hs::solve (src/driver/solver.cpp:523-694). Here the same call takes the matrices as
 * CSR arrays in host memory and a config JSON document, runs the solve in this engine (its driver
 * is the hsolve_amd Python package over libhsg.so's GPU kernels) and hands back the eigenvalues,
 * the eigenvectors and the two JSON documents through an opaque result object.
 *
 * Library: libhsg_solve.so (next to libhsg.so in the hsolve_amd package directory). It runs the
 * driver in the process's Python interpreter: inside a Python process (ctypes, cffi) it uses the
 * running one; from a C or C++ program it starts one on the first call, and that program links
 * libpython3 itself (-lpython3.10), since this library does not carry a second copy of the
 * interpreter. The hsolve_amd package is found next to this library. Calls are serialised (one
 * solve at a time per process, as module.cpp:14-16).
 */
#ifndef HSG_SOLVE_H
#define HSG_SOLVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: as in hsg.h (0 = success; hs::Error::Code order plus one,
 * include/hsolve/error.hpp:15-25). */
enum {
  HSG_SOLVE_OK = 0,
  HSG_SOLVE_INVALID_INPUT = 1,
  HSG_SOLVE_IO = 2,
  HSG_SOLVE_NOT_HERMITIAN = 3,
  HSG_SOLVE_UNSUPPORTED = 4,
  HSG_SOLVE_FACTORIZATION_FAILED = 5,
  HSG_SOLVE_SINGULAR = 6,
  HSG_SOLVE_OUT_OF_MEMORY_BUDGET = 7,
  HSG_SOLVE_NOT_CONVERGED = 8,
  HSG_SOLVE_INTERNAL = 9
};

/* A square sparse matrix in CSR form, host memory, zero-based: row i holds the entries
 * indices[indptr[i] .. indptr[i+1]) with values at the same positions (indptr[0] = 0). values
 * holds nnz doubles, or 2 * nnz doubles (real, imaginary interleaved) when is_complex is nonzero.
 * As with the reference's matrix argument (module.cpp matrix_arg), `storage` says which
 * triangles the arrays hold. The arrays are read during the call only. */
typedef struct hsg_matrix {
  int64_t n;
  const int64_t* indptr;  /* n + 1 */
  const int32_t* indices; /* nnz */
  const double* values;   /* nnz, or 2 * nnz when is_complex */
  int32_t is_complex;
  int32_t reserved;
} hsg_matrix;

typedef struct hsg_result hsg_result;

/* Solves A x = lambda B x (b = NULL: A x = lambda x) for Hermitian A and Hermitian positive
 * definite B. config_json: a config document (the reference's config.json keys; NULL or "" =
 * the defaults); verbose: 0, 1 or 2; storage: "auto", "full", "upper" or "lower" (NULL =
 * "auto"). On success *out receives the result (free it with hsg_result_free); a solve that
 * ends without meeting its tolerance still succeeds (the report's "success" is false). On error
 * *out is NULL and hsg_solve_last_error_message() holds the message. */
int hsg_solve(const hsg_matrix* a, const hsg_matrix* b, const char* config_json, int verbose, const char* storage,
              hsg_result** out);

/* The result: n, the number m of eigenpairs, whether the vectors are complex (complex input),
 * the eigenvalues (m, ascending as reported) and the eigenvectors (n x m, column-major; for
 * complex input 2 doubles per entry, real and imaginary interleaved), the report and the
 * effective config as JSON text (report.json / config.json of the reference). The pointers stay
 * valid until hsg_result_free. */
int64_t hsg_result_order(const hsg_result* s);
int64_t hsg_result_count(const hsg_result* s);
int32_t hsg_result_is_complex(const hsg_result* s);
const double* hsg_result_eigenvalues(const hsg_result* s);
const double* hsg_result_eigenvectors(const hsg_result* s);
const char* hsg_result_report_json(const hsg_result* s);
const char* hsg_result_config_json(const hsg_result* s);
void hsg_result_free(hsg_result* s);

/* Validates a config document and completes it with the defaults (the reference's
 * _hsolve.parse_config, module.cpp:466-467: parse_config then config_to_json). Returns the
 * complete document as JSON text, valid until the next call on this thread, or NULL on invalid
 * input (the message in hsg_solve_last_error_message()). */
const char* hsg_parse_config(const char* config_json);

/* Message of the last failing hsg_solve or hsg_parse_config on this thread (empty if none). */
const char* hsg_solve_last_error_message(void);

#ifdef __cplusplus
}
#endif

#endif /* HSG_SOLVE_H */
