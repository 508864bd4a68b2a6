/* A C host of hsg_solve (include/hsg_solve.h): the 1D Dirichlet Laplacian pencil
 * A = tridiag(-1, 2, -1), B = 2 I of order n (eigenvalues (1 - cos(j pi / (n + 1))), j = 1..n),
 * the k eigenvalues nearest sigma. Prints the return code, then one eigenvalue per line and the
 * largest residual |A x - lambda B x|_inf / |x|_inf. Exit status 0 on success, 2 otherwise.
 *
 *   solve_demo n sigma k
 */
#include <math.h>
#include <string.h>
#include <stdio.h>
#include <stdlib.h>

#include "hsg_solve.h"

int main(int argc, char** argv) {
  const int64_t n = argc > 1 ? atoll(argv[1]) : 2000;
  const double sigma = argc > 2 ? atof(argv[2]) : 1.0;
  const int k = argc > 3 ? atoi(argv[3]) : 4;
  int64_t* ap = malloc(sizeof(int64_t) * (n + 1));
  int32_t* ai = malloc(sizeof(int32_t) * 3 * n);
  double* av = malloc(sizeof(double) * 3 * n);
  int64_t* bp = malloc(sizeof(int64_t) * (n + 1));
  int32_t* bi = malloc(sizeof(int32_t) * n);
  double* bv = malloc(sizeof(double) * n);
  int64_t q = 0;
  for (int64_t i = 0; i < n; ++i) {
    ap[i] = q;
    if (i > 0) { ai[q] = (int32_t)(i - 1); av[q++] = -1.0; }
    ai[q] = (int32_t)i; av[q++] = 2.0;
    if (i + 1 < n) { ai[q] = (int32_t)(i + 1); av[q++] = -1.0; }
    bp[i] = i; bi[i] = (int32_t)i; bv[i] = 2.0;
  }
  ap[n] = q;
  bp[n] = n;
  hsg_matrix a = {n, ap, ai, av, 0, 0}, b = {n, bp, bi, bv, 0, 0};
  char config[256];
  snprintf(config, sizeof config, "{\"target\": {\"mode\": \"nearest\", \"sigma\": %.17g, \"k\": %d}}", sigma, k);
  hsg_result* s = NULL;
  const int rc = hsg_solve(&a, &b, config, 0, "full", &s);
  printf("rc %d\n", rc);
  if (rc != HSG_SOLVE_OK) {
    printf("error %s\n", hsg_solve_last_error_message());
    return 2;
  }
  const int64_t m = hsg_result_count(s);
  const double* w = hsg_result_eigenvalues(s);
  const double* x = hsg_result_eigenvectors(s);
  double worst = 0.0;
  for (int64_t j = 0; j < m; ++j) {
    const double* v = x + j * n;
    double r = 0.0, nv = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      double ax = 2.0 * v[i];
      if (i > 0) ax -= v[i - 1];
      if (i + 1 < n) ax -= v[i + 1];
      r = fmax(r, fabs(ax - w[j] * 2.0 * v[i]));
      nv = fmax(nv, fabs(v[i]));
    }
    worst = fmax(worst, r / nv);
    printf("lambda %.17g\n", w[j]);
  }
  printf("residual %.3e\n", worst);
  printf("report_bytes %zu\n", strlen(hsg_result_report_json(s)));
  hsg_result_free(s);
  free(ap); free(ai); free(av); free(bp); free(bi); free(bv);
  return 0;
}
