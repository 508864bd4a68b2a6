// CPU multifrontal LDL^T of a sparse symmetric matrix (oracle / CPU baseline; test infrastructure,
// never linked into the product library).
//
// The CPU counterpart of the reference's factorization path: MUMPS' multifrontal LDL^T (SYM = 2,
// src/factor/mumps_factorization.cpp:299-381) with Bunch-Kaufman 1x1 / 2x2 pivots, the inertia
// read off D (INFOG(12)), and the multi-right-hand-side solve (JOB = 3, :458-527). It runs over a
// given supernodal symbolic analysis (fronts: first column, pivot count, rows, parent; new
// numbering, order new -> old), factorizes fronts level by level of the assembly tree with
// OpenMP (whole fronts per thread on the wide lower levels, the GEMM/TRSM of one front split
// over threads on the narrow top levels) and calls a BLAS for the dense kernels (dgemm / dtrsm
// passed in as Fortran-ABI function pointers, e.g. the OpenBLAS of the Python environment).
//
// Per front (order m, ncol pivot columns, dense lower triangle in a column-major m x m array):
// original entries assembled, children's contribution blocks added, then panels of nb columns:
// Bunch-Kaufman on the nb x nb diagonal block with pivot candidates inside the block (the same
// restriction the GPU engine uses), X = A21 L11^-T by dtrsm, L21 = X D^-1, trailing update
// A22 -= X L21^T (lower part, dgemm by column blocks). The contribution block (rows / columns
// ncol..m) is handed to the parent and freed once the parent has assembled it.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

typedef void (*dgemm_t)(const char*, const char*, const int*, const int*, const int*, const double*, const double*,
                        const int*, const double*, const int*, const double*, double*, const int*);
typedef void (*dtrsm_t)(const char*, const char*, const char*, const char*, const int*, const int*, const double*,
                        const double*, const int*, double*, const int*);
dgemm_t g_dgemm = nullptr;
dtrsm_t g_dtrsm = nullptr;

constexpr double kAlpha = 0.6403882032022076;  // Bunch-Kaufman (1 + sqrt(17)) / 8
constexpr int kNb = 64;                         // panel width
constexpr int kChunk = 128;                     // column chunk of a split trailing update

void gemm_nt(int mr, int nc, int k, double alpha, const double* a, int lda, const double* b, int ldb, double beta,
             double* c, int ldc) {
  if (mr <= 0 || nc <= 0 || k <= 0) return;
  const char n = 'N', t = 'T';
  g_dgemm(&n, &t, &mr, &nc, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
}

void gemm_nn(int mr, int nc, int k, double alpha, const double* a, int lda, const double* b, int ldb, double beta,
             double* c, int ldc) {
  if (mr <= 0 || nc <= 0 || k <= 0) return;
  const char n = 'N';
  g_dgemm(&n, &n, &mr, &nc, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
}

void gemm_tn(int mr, int nc, int k, double alpha, const double* a, int lda, const double* b, int ldb, double beta,
             double* c, int ldc) {
  if (mr <= 0 || nc <= 0 || k <= 0) return;
  const char n = 'N', t = 'T';
  g_dgemm(&t, &n, &mr, &nc, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
}

// B := B op(L)^-1 or op(L)^-1 B for unit lower L (side 'R'/'L', trans 'N'/'T')
void trsm_unit_lower(char side, char trans, int mr, int nc, const double* l, int ldl, double* b, int ldb) {
  if (mr <= 0 || nc <= 0) return;
  const char lo = 'L', u = 'U';
  const double one = 1.0;
  g_dtrsm(&side, &lo, &trans, &u, &mr, &nc, &one, l, &ldl, b, &ldb);
}

struct Analysis {
  int n = 0, nsn = 0;
  std::vector<int> first, ncol, m, parent, order;  // order: new -> old
  std::vector<int64_t> rowoff;
  std::vector<int> rows;                            // new numbering, ascending per front
  std::vector<int> relmap;                          // per front row (CB part): local row in the parent
  std::vector<std::vector<int>> children, levels;
  std::vector<int64_t> panel_off;                   // m x ncol panel per front
  // assembly of the original pattern (CSR, old numbering): entry -> (front, local col, local row), or -1
  std::vector<int64_t> asm_front_pos;               // front-local index col * m + row
  std::vector<int> asm_front;                       // the front
  int64_t nnz = 0;
  // numeric
  std::vector<double> L;                            // panels, column-major ld m; D on the diagonal
  std::vector<double> doff;                         // per pivot position: D(k+1, k) of a 2x2 pivot at k
  std::vector<int8_t> ptype;                        // 0: 1x1, 1: first of 2x2, 2: second, 3: zero pivot
  std::vector<int> perm;                            // per pivot position first + i: local front row
  int64_t neg = 0, zero = 0, pos = 0, two = 0;
  bool factored = false;
};

std::string g_error;

// Bunch-Kaufman on the b x b diagonal block at (k0, k0) of the front F (ld m), pivot candidates
// restricted to the block. Symmetric interchanges are applied to the block's lower triangle, to
// the rows of the previous panels (columns < k0) and to the columns of the rows below the block.
// On exit the block holds unit L11 (2x2 pivots: identity diagonal block, D(k+1,k) in doff) and D on
// the diagonal. lperm[i]: local front row that ends at block position i.
void bk_block(double* F, int m, int k0, int b, int* lperm, double* doffv, int8_t* pt, int64_t cnt[4]) {
  auto at = [&](int i, int j) -> double& { return F[(int64_t)j * m + i]; };  // i >= j
  auto sym = [&](int i, int j) -> double { return i >= j ? at(i, j) : at(j, i); };
  auto swap_sym = [&](int p, int q) {  // p < q, global front positions
    if (p == q) return;
    for (int j = 0; j < p; ++j) std::swap(at(p, j), at(q, j));           // rows of earlier columns
    for (int j = p + 1; j < q; ++j) std::swap(at(j, p), at(q, j));       // between
    for (int i = q + 1; i < m; ++i) std::swap(at(i, p), at(i, q));       // below (incl. rows past the block)
    std::swap(at(p, p), at(q, q));
  };
  int k = 0;
  while (k < b) {
    const int gk = k0 + k;
    const double absakk = std::fabs(at(gk, gk));
    double colmax = 0.0;
    int imax = -1;
    for (int i = k + 1; i < b; ++i) {
      const double v = std::fabs(at(k0 + i, gk));
      if (v > colmax) {
        colmax = v;
        imax = i;
      }
    }
    int kind = 1, kp = k;
    if (!(absakk >= kAlpha * colmax) && imax > k) {
      double rowmax = 0.0;
      for (int j = k; j < b; ++j)
        if (j != imax) rowmax = std::max(rowmax, std::fabs(sym(k0 + imax, k0 + j)));
      if (absakk * rowmax >= kAlpha * colmax * colmax) {
        kind = 1;
        kp = k;
      } else if (std::fabs(at(k0 + imax, k0 + imax)) >= kAlpha * rowmax) {
        kind = 1;
        kp = imax;
      } else {
        kind = 2;
        kp = imax;
      }
    }
    if (kind == 2 && k + 1 >= b) kind = 1, kp = k;  // no room for a 2x2 pivot at the block end
    const int kk = kind == 1 ? k : k + 1;
    if (kp != kk) {
      swap_sym(k0 + kk, k0 + kp);
      std::swap(lperm[kk], lperm[kp]);
    }
    if (kind == 1) {
      const double d = at(gk, gk);
      pt[k] = 0;
      if (d == 0.0) {
        pt[k] = 3;
        ++cnt[1];
        // a zero pivot with a zero column (the block is rank deficient): nothing to eliminate
        for (int i = k + 1; i < m - k0; ++i) at(k0 + i, gk) = 0.0;
      } else {
        ++cnt[d < 0 ? 0 : 2];
        const double inv = 1.0 / d;
        for (int j = k + 1; j < b; ++j) {  // rank-1 update of the rest of the block
          const double lj = at(k0 + j, gk) * inv;
          for (int i = j; i < b; ++i) at(k0 + i, k0 + j) -= at(k0 + i, gk) * lj;
        }
        for (int i = k + 1; i < b; ++i) at(k0 + i, gk) *= inv;
      }
      doffv[k] = 0.0;
      k += 1;
    } else {
      const double a = at(gk, gk), c = at(gk + 1, gk + 1), o = at(gk + 1, gk);
      const double det = a * c - o * o;
      // eigenvalues of [[a, o], [o, c]]: signs from det and trace
      if (det < 0) {
        ++cnt[0];
        ++cnt[2];
      } else if (a + c < 0) {
        cnt[0] += 2;
      } else {
        cnt[2] += 2;
      }
      ++cnt[3];
      const double ia = c / det, ic = a / det, io = -o / det;  // D^-1
      for (int i = k + 2; i < b; ++i) {
        const double x0 = at(k0 + i, gk), x1 = at(k0 + i, gk + 1);
        const double l0 = x0 * ia + x1 * io, l1 = x0 * io + x1 * ic;
        for (int j = k + 2; j <= i; ++j) {
          const double y0 = at(k0 + j, gk), y1 = at(k0 + j, gk + 1);
          at(k0 + i, k0 + j) -= l0 * y0 + l1 * y1;
        }
      }
      for (int i = k + 2; i < b; ++i) {
        const double x0 = at(k0 + i, gk), x1 = at(k0 + i, gk + 1);
        at(k0 + i, gk) = x0 * ia + x1 * io;
        at(k0 + i, gk + 1) = x0 * io + x1 * ic;
      }
      pt[k] = 1;
      pt[k + 1] = 2;
      doffv[k] = o;
      doffv[k + 1] = 0.0;
      at(gk + 1, gk) = 0.0;  // unit L11 has an identity 2x2 block here
      k += 2;
    }
  }
}

// X := X D^-1 for the rows r0..m of the block columns (1x1 and 2x2 blocks of D).
void scale_by_dinv(double* F, int m, int k0, int b, int r0, const double* doffv, const int8_t* pt, double* X, int ldx) {
  for (int k = 0; k < b;) {
    if (pt[k] == 1) {
      const double a = F[(int64_t)(k0 + k) * m + k0 + k], c = F[(int64_t)(k0 + k + 1) * m + k0 + k + 1], o = doffv[k];
      const double det = a * c - o * o, ia = c / det, ic = a / det, io = -o / det;
      double* x0 = X + (int64_t)k * ldx;
      double* x1 = X + (int64_t)(k + 1) * ldx;
      for (int i = 0; i < m - r0; ++i) {
        const double u = x0[i], v = x1[i];
        x0[i] = u * ia + v * io;
        x1[i] = u * io + v * ic;
      }
      k += 2;
    } else {
      const double d = F[(int64_t)(k0 + k) * m + k0 + k];
      double* x = X + (int64_t)k * ldx;
      const double inv = pt[k] == 3 ? 0.0 : 1.0 / d;
      for (int i = 0; i < m - r0; ++i) x[i] *= inv;
      k += 1;
    }
  }
}

void factor_front(Analysis& A, int s, std::vector<std::unique_ptr<double[]>>& cbs, const double* vals, bool split,
                  int64_t cnt[4]) {
  const int m = A.m[s], nc = A.ncol[s], first = A.first[s];
  std::unique_ptr<double[]> Fp(new double[(size_t)m * m]);
  double* F = Fp.get();
  (void)vals;
  // only the lower triangle is referenced: zero it column by column (the pivot columns are
  // overwritten by the assembled panel below)
  for (int j = nc; j < m; ++j) std::memset(F + (int64_t)j * m + j, 0, sizeof(double) * (m - j));
  // original entries: the precomputed (front, position) of every pattern entry
  // (assembled by the caller into the panel array, copied here)
  const double* P = A.L.data() + A.panel_off[s];
  for (int j = 0; j < nc; ++j) std::memcpy(F + (int64_t)j * m + j, P + (int64_t)j * m + j, sizeof(double) * (m - j));
  // children's contribution blocks
  for (int c : A.children[s]) {
    const int cm = A.m[c], cn = A.ncol[c], cb = cm - cn;
    const double* C = cbs[c].get();
    const int* rm = A.relmap.data() + A.rowoff[c] + cn;
    for (int b = 0; b < cb; ++b) {
      double* dst = F + (int64_t)rm[b] * m;
      const double* src = C + (int64_t)b * cb;
      for (int a = b; a < cb; ++a) dst[rm[a]] += src[a];
    }
    cbs[c].reset();
  }
  std::vector<int> lperm(nc);
  for (int i = 0; i < nc; ++i) lperm[i] = i;
  for (int k0 = 0; k0 < nc; k0 += kNb) {
    const int b = std::min(kNb, nc - k0);
    bk_block(F, m, k0, b, lperm.data() + k0, A.doff.data() + first + k0, A.ptype.data() + first + k0, cnt);
    const int r0 = k0 + b;  // rows below the block
    const int nr = m - r0;
    if (nr <= 0) continue;
    double* A21 = F + (int64_t)k0 * m + r0;  // nr x b, ld m
    const int chunks = (nr + kChunk - 1) / kChunk;
    // X = A21 L11^-T (X = L21 D), by row chunks
    auto solve_rows = [&](int q) {
      const int c0 = q * kChunk, cw = std::min(kChunk, nr - c0);
      trsm_unit_lower('R', 'T', cw, b, F + (int64_t)k0 * m + k0, m, A21 + c0, m);
    };
    if (split && chunks > 1) {
#pragma omp taskloop grainsize(1)
      for (int q = 0; q < chunks; ++q) solve_rows(q);
    } else {
      for (int q = 0; q < chunks; ++q) solve_rows(q);
    }
    std::vector<double> X((size_t)nr * b);
    for (int j = 0; j < b; ++j) std::memcpy(X.data() + (size_t)j * nr, A21 + (int64_t)j * m, sizeof(double) * nr);
    scale_by_dinv(F, m, k0, b, r0, A.doff.data() + first + k0, A.ptype.data() + first + k0, A21, m);  // L21
    // A22 -= X L21^T, lower triangle, by column chunks
    auto chunk = [&](int q) {
      const int c0 = q * kChunk, cw = std::min(kChunk, nr - c0);
      gemm_nt(nr - c0, cw, b, -1.0, X.data() + c0, nr, A21 + c0, m, 1.0, F + (int64_t)(r0 + c0) * m + r0 + c0, m);
    };
    if (split && chunks > 1) {
#pragma omp taskloop grainsize(1)
      for (int q = 0; q < chunks; ++q) chunk(q);
    } else {
      for (int q = 0; q < chunks; ++q) chunk(q);
    }
  }
  for (int i = 0; i < nc; ++i) A.perm[first + i] = lperm[i];
  // panel (pivot columns in pivot order, rows: first nc in pivot order, then the CB rows)
  double* Pw = A.L.data() + A.panel_off[s];
  for (int j = 0; j < nc; ++j) std::memcpy(Pw + (int64_t)j * m, F + (int64_t)j * m, sizeof(double) * m);
  const int cb = m - nc;
  if (cb > 0) {
    std::unique_ptr<double[]> C(new double[(size_t)cb * cb]);
    for (int b = 0; b < cb; ++b)
      std::memcpy(C.get() + (int64_t)b * cb + b, F + (int64_t)(nc + b) * m + nc + b, sizeof(double) * (cb - b));
    cbs[s] = std::move(C);
  }
}

}  // namespace

extern "C" {

const char* mf_last_error() { return g_error.c_str(); }

void mf_set_blas(void* dgemm, void* dtrsm) {
  g_dgemm = reinterpret_cast<dgemm_t>(dgemm);
  g_dtrsm = reinterpret_cast<dtrsm_t>(dtrsm);
}

// Fronts of a supernodal analysis (new numbering) and the CSR pattern (old numbering, both
// triangles) whose values mf_factorize assembles. Returns a handle or NULL (mf_last_error).
void* mf_analyze(int n, int nsn, const int* first, const int* ncol, const int* m, const int* parent,
                 const int64_t* rowoff, const int* rows, const int* order, const int64_t* rowptr, const int* colind) {
  try {
    auto A = std::make_unique<Analysis>();
    A->n = n;
    A->nsn = nsn;
    A->first.assign(first, first + nsn);
    A->ncol.assign(ncol, ncol + nsn);
    A->m.assign(m, m + nsn);
    A->parent.assign(parent, parent + nsn);
    A->rowoff.assign(rowoff, rowoff + nsn + 1);
    A->rows.assign(rows, rows + rowoff[nsn]);
    A->order.assign(order, order + n);
    std::vector<int> iorder(n), col_sn(n);
    for (int i = 0; i < n; ++i) iorder[order[i]] = i;
    A->children.assign(nsn, {});
    std::vector<int> level(nsn, 0);
    for (int s = 0; s < nsn; ++s) {
      for (int j = 0; j < ncol[s]; ++j) col_sn[first[s] + j] = s;
      if (parent[s] >= 0) A->children[parent[s]].push_back(s);
    }
    int nlev = 0;
    for (int s = 0; s < nsn; ++s) {  // children precede parents in the supernodal numbering
      for (int c : A->children[s]) level[s] = std::max(level[s], level[c] + 1);
      nlev = std::max(nlev, level[s] + 1);
    }
    A->levels.assign(nlev, {});
    for (int s = 0; s < nsn; ++s) A->levels[level[s]].push_back(s);
    A->relmap.assign(A->rows.size(), -1);
    for (int c = 0; c < nsn; ++c) {
      const int p = parent[c];
      if (p < 0) continue;
      const int* pr = A->rows.data() + rowoff[p];
      int q = 0;
      for (int64_t k = rowoff[c] + ncol[c]; k < rowoff[c + 1]; ++k) {
        while (q < m[p] && pr[q] < A->rows[k]) ++q;
        if (q >= m[p] || pr[q] != A->rows[k]) throw std::runtime_error("child row missing in parent");
        A->relmap[k] = q;
      }
    }
    A->panel_off.assign(nsn + 1, 0);
    for (int s = 0; s < nsn; ++s) A->panel_off[s + 1] = A->panel_off[s] + (int64_t)m[s] * ncol[s];
    A->nnz = rowptr[n];
    A->asm_front_pos.assign(A->nnz, -1);
    for (int r = 0; r < n; ++r)
      for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e) {
        const int i = iorder[r], j = iorder[colind[e]];
        if (i < j) continue;
        const int s = col_sn[j];
        const int* fr = A->rows.data() + rowoff[s];
        const int loc = (int)(std::lower_bound(fr, fr + m[s], i) - fr);
        if (loc >= m[s] || fr[loc] != i) throw std::runtime_error("pattern entry outside its front");
        A->asm_front_pos[e] = A->panel_off[s] + (int64_t)(j - first[s]) * m[s] + loc;
      }
    return A.release();
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

// Numeric factorization of K (values aligned with the CSR pattern of mf_analyze). inertia[3]:
// negative, zero, positive; returns 0 or -1 (mf_last_error).
int mf_factorize(void* h, const double* values, int64_t* inertia) {
  try {
    if (!g_dgemm || !g_dtrsm) throw std::runtime_error("BLAS not set (mf_set_blas)");
    Analysis& A = *static_cast<Analysis*>(h);
    A.L.assign(A.panel_off[A.nsn], 0.0);
    A.doff.assign(A.n, 0.0);
    A.ptype.assign(A.n, 0);
    A.perm.assign(A.n, 0);
    for (int64_t e = 0; e < A.nnz; ++e)
      if (A.asm_front_pos[e] >= 0) A.L[A.asm_front_pos[e]] = values[e];
    std::vector<std::unique_ptr<double[]>> cbs(A.nsn);
    int64_t total[4] = {0, 0, 0, 0};
    // tree parallelism: a task per front after its children's tasks; fronts of at least
    // kSplitOrder rows also split their row solves and updates into tasks (taskloop)
    constexpr int kSplitOrder = 512;
    std::vector<int> roots;
    for (int s = 0; s < A.nsn; ++s)
      if (A.parent[s] < 0) roots.push_back(s);
    std::function<void(int)> run = [&](int s) {
      for (int c : A.children[s]) {
#pragma omp task firstprivate(c) shared(run)
        run(c);
      }
#pragma omp taskwait
      int64_t loc[4] = {0, 0, 0, 0};
      factor_front(A, s, cbs, values, A.m[s] >= kSplitOrder, loc);
      for (int t = 0; t < 4; ++t) {
#pragma omp atomic
        total[t] += loc[t];
      }
    };
#pragma omp parallel
#pragma omp single
    {
      for (int r : roots) {
#pragma omp task firstprivate(r) shared(run)
        run(r);
      }
#pragma omp taskwait
    }
    A.neg = total[0];
    A.zero = total[1];
    A.pos = total[2];
    A.two = total[3];
    A.factored = true;
    inertia[0] = A.neg;
    inertia[1] = A.zero;
    inertia[2] = A.pos;
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

// In-place solve K X = X for p columns (column-major, leading dimension ldx, old numbering).
// Forward sweep bottom-up and backward sweep top-down over the assembly tree as OpenMP tasks (a
// front after its children, resp. before them); the products of fronts of at least 512 rows are
// split over row chunks. Each front's update rows go to its own buffer (U), gathered by the parent.
int mf_solve(void* h, double* x, int p, int64_t ldx) {
  try {
    Analysis& A = *static_cast<Analysis*>(h);
    if (!A.factored) throw std::runtime_error("solve without a factorization");
    const int n = A.n, nsn = A.nsn;
    std::vector<double> Y((size_t)n * p), Xn((size_t)n * p);  // new numbering, ld n
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < p; ++c) Y[(size_t)c * n + i] = x[(int64_t)c * ldx + A.order[i]];
    // workspace offsets: U (cb x p) and Z (ncol x p) per front
    std::vector<int64_t> uoff(nsn + 1, 0), zoff(nsn + 1, 0);
    for (int s = 0; s < nsn; ++s) {
      uoff[s + 1] = uoff[s] + (int64_t)(A.m[s] - A.ncol[s]) * p;
      zoff[s + 1] = zoff[s] + (int64_t)A.ncol[s] * p;
    }
    std::vector<double> U(uoff[nsn]), Z(zoff[nsn]);
    constexpr int kSplit = 512;
    auto rows_split = [&](int rows, const std::function<void(int, int)>& f, bool split) {
      const int chunks = (rows + kChunk - 1) / kChunk;
      if (split && chunks > 1) {
#pragma omp taskloop grainsize(1)
        for (int q = 0; q < chunks; ++q) f(q * kChunk, std::min(kChunk, rows - q * kChunk));
      } else if (rows > 0) {
        f(0, rows);
      }
    };
    auto forward = [&](int s) {
      const int m = A.m[s], nc = A.ncol[s], first = A.first[s], cb = m - nc;
      const bool split = m >= kSplit;
      double* z = Z.data() + zoff[s];
      double* u = U.data() + uoff[s];
      const int* perm = A.perm.data() + first;
      // z: own rows in pivot order; u: update rows; children's update rows added to both
      std::vector<int> inv(nc);
      for (int i = 0; i < nc; ++i) inv[perm[i]] = i;
      for (int c = 0; c < p; ++c) {
        for (int i = 0; i < nc; ++i) z[(size_t)c * nc + inv[i]] = Y[(size_t)c * n + first + i];
        std::fill(u + (size_t)c * cb, u + (size_t)(c + 1) * cb, 0.0);
      }
      for (int ch : A.children[s]) {
        const int ccb = A.m[ch] - A.ncol[ch];
        const int* rm = A.relmap.data() + A.rowoff[ch] + A.ncol[ch];
        const double* uc = U.data() + uoff[ch];
        for (int c = 0; c < p; ++c)
          for (int a = 0; a < ccb; ++a) {
            const int r = rm[a];
            if (r < nc) z[(size_t)c * nc + inv[r]] += uc[(size_t)c * ccb + a];
            else u[(size_t)c * cb + r - nc] += uc[(size_t)c * ccb + a];
          }
      }
      const double* P = A.L.data() + A.panel_off[s];
      trsm_unit_lower('L', 'N', nc, p, P, m, z, nc);
      rows_split(cb, [&](int r0, int rw) { gemm_nn(rw, p, nc, -1.0, P + nc + r0, m, z, nc, 1.0, u + r0, cb); }, split);
      for (int k = 0; k < nc;) {  // z := D^-1 z
        const int8_t t = A.ptype[first + k];
        if (t == 1) {
          const double a = P[(int64_t)k * m + k], cc = P[(int64_t)(k + 1) * m + k + 1], o = A.doff[first + k];
          const double det = a * cc - o * o, ia = cc / det, ic = a / det, io = -o / det;
          for (int c = 0; c < p; ++c) {
            double* zz = z + (size_t)c * nc;
            const double u0 = zz[k], u1 = zz[k + 1];
            zz[k] = u0 * ia + u1 * io;
            zz[k + 1] = u0 * io + u1 * ic;
          }
          k += 2;
        } else {
          const double inv_d = t == 3 ? 0.0 : 1.0 / P[(int64_t)k * m + k];
          for (int c = 0; c < p; ++c) z[(size_t)c * nc + k] *= inv_d;
          k += 1;
        }
      }
    };
    auto backward = [&](int s) {
      const int m = A.m[s], nc = A.ncol[s], first = A.first[s], cb = m - nc;
      const bool split = nc >= kSplit;
      double* z = Z.data() + zoff[s];
      const double* P = A.L.data() + A.panel_off[s];
      if (cb > 0) {
        std::vector<double> xc((size_t)cb * p);
        const int* fr = A.rows.data() + A.rowoff[s] + nc;
        for (int c = 0; c < p; ++c)
          for (int a = 0; a < cb; ++a) xc[(size_t)c * cb + a] = Xn[(size_t)c * n + fr[a]];
        rows_split(nc, [&](int r0, int rw) {
          gemm_tn(rw, p, cb, -1.0, P + (int64_t)r0 * m + nc, m, xc.data(), cb, 1.0, z + r0, nc);
        }, split);
      }
      trsm_unit_lower('L', 'T', nc, p, P, m, z, nc);
      const int* perm = A.perm.data() + first;
      for (int c = 0; c < p; ++c)
        for (int i = 0; i < nc; ++i) Xn[(size_t)c * n + first + perm[i]] = z[(size_t)c * nc + i];
    };
    std::function<void(int)> run_fwd = [&](int s) {
      for (int c : A.children[s]) {
#pragma omp task firstprivate(c) shared(run_fwd)
        run_fwd(c);
      }
#pragma omp taskwait
      forward(s);
    };
    std::function<void(int)> run_bwd = [&](int s) {
      backward(s);
      for (int c : A.children[s]) {
#pragma omp task firstprivate(c) shared(run_bwd)
        run_bwd(c);
      }
#pragma omp taskwait
    };
    std::vector<int> roots;
    for (int s = 0; s < nsn; ++s)
      if (A.parent[s] < 0) roots.push_back(s);
#pragma omp parallel
#pragma omp single
    {
      for (int r : roots) {
#pragma omp task firstprivate(r) shared(run_fwd)
        run_fwd(r);
      }
#pragma omp taskwait
      for (int r : roots) {
#pragma omp task firstprivate(r) shared(run_bwd)
        run_bwd(r);
      }
#pragma omp taskwait
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < p; ++c) x[(int64_t)c * ldx + A.order[i]] = Xn[(size_t)c * n + i];
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

// Dense block operations of the Krylov layer on (rows = vectors) blocks: V is k vectors of length n
// (column-major n x k, ld n), W p vectors. Row chunks run on the OpenMP threads, each calling a
// single-threaded dgemm; the Gram partial sums of the chunks are added in chunk order.
void mf_gram(int64_t n, const double* v, int k, const double* w, int p, const double* weight, double* c) {
  constexpr int64_t kRows = 8192;
  const int64_t chunks = (n + kRows - 1) / kRows;
  std::vector<double> part((size_t)chunks * k * p, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < chunks; ++q) {
    const int64_t r0 = q * kRows;
    const int rows = (int)std::min<int64_t>(kRows, n - r0);
    std::vector<double> ww;
    const double* wp = w + r0;
    int ldw = (int)n;
    if (weight) {
      ww.resize((size_t)rows * p);
      for (int c2 = 0; c2 < p; ++c2)
        for (int i = 0; i < rows; ++i) ww[(size_t)c2 * rows + i] = w[(int64_t)c2 * n + r0 + i] * weight[r0 + i];
      wp = ww.data();
      ldw = rows;
    }
    const int ldv = (int)n;
    gemm_tn(k, p, rows, 1.0, v + r0, ldv, wp, ldw, 0.0, part.data() + (size_t)q * k * p, k);
  }
  for (int64_t e = 0; e < (int64_t)k * p; ++e) {
    double acc = 0.0;
    for (int64_t q = 0; q < chunks; ++q) acc += part[(size_t)q * k * p + e];
    c[e] = acc;
  }
}

// W = beta W + alpha V C (C: k x p column-major), by row chunks.
void mf_update(int64_t n, const double* v, int k, const double* cmat, double* w, int p, double alpha, double beta) {
  constexpr int64_t kRows = 8192;
  const int64_t chunks = (n + kRows - 1) / kRows;
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < chunks; ++q) {
    const int64_t r0 = q * kRows;
    const int rows = (int)std::min<int64_t>(kRows, n - r0);
    if (k == 0) {
      for (int c2 = 0; c2 < p; ++c2)
        for (int i = 0; i < rows; ++i) w[(int64_t)c2 * n + r0 + i] *= beta;
      continue;
    }
    const int ld = (int)n;
    gemm_nn(rows, p, k, alpha, v + r0, ld, cmat, k, beta, w + r0, ld);
  }
}

double mf_factor_entries(void* h) {
  const Analysis& A = *static_cast<Analysis*>(h);
  double e = 0;
  for (int s = 0; s < A.nsn; ++s) e += (double)A.m[s] * A.ncol[s] - 0.5 * (double)A.ncol[s] * (A.ncol[s] - 1);
  return e;
}

void mf_free(void* h) { delete static_cast<Analysis*>(h); }

int mf_threads() { return omp_get_max_threads(); }

}  // extern "C"
