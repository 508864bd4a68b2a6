// Multifrontal LDL^T with Bunch-Kaufman pivoting inside fronts, on CDNA4 (gfx950).
//
// Replaces the numeric factorization the reference delegates to MUMPS (SYM = 2, JOB = 2;
// src/factor/mumps_factorization.cpp:299-356) and the inertia it reads from INFOG(12)/(28)
// (:358-381). Fronts are processed level by level of the assembly tree (children before parents):
//   * small fronts (order m <= small_max): one workgroup per front, the whole front in LDS:
//     load the panel (original entries were scattered there beforehand), extend-add the children's
//     contribution blocks, factor the pivot columns with Bunch-Kaufman 1x1/2x2 pivots restricted
//     to the front's fully-summed block, write L, D and the contribution block back;
//   * large fronts: extend-add by column ranges (each parent column owned by one workgroup, children
//     in order: deterministic), then per panel of nb pivot columns a one-workgroup Bunch-Kaufman
//     factorization of the nb x nb diagonal block, a multi-workgroup panel solve that forms
//     W = A21 L11^-T and L21 = W D^-1, and a trailing GEMM update of the remaining pivot columns;
//     finally one GEMM contribution-block update CB -= W_cb L_cb^T over all pivot columns.
// Inertia: signs of D (2x2 blocks by their eigenvalues), integer counters. Null pivots: a pivot
// block with an eigenvalue of magnitude <= null_tol * max|K_ij| is flagged (the reference's
// CNTL(3) rule, include/hsolve/factor.hpp:79-115), counted as zero and removed from the
// elimination (like MUMPS' treatment of null pivots).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"
#include "front_common.hpp"

namespace hsg {

constexpr int kThreads = 256;

// ------------------------------------------------------------------------------ small helpers

template <class T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Original entries of K = alpha A + beta B into the panels; the largest |K_ij| is reduced into
// the counter slot kCntKnorm (bit pattern of a non-negative double: integer max is value max).
template <class T>
__global__ void k_assemble_original(int64_t nnz, const int64_t* __restrict__ dst, const int64_t* __restrict__ ia,
                                    const int64_t* __restrict__ ib, const double* __restrict__ av,
                                    const double* __restrict__ bv, double alpha, double beta, T* __restrict__ L,
                                    unsigned long long* cnt) {
  double local_max = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = dst[e];
    if (d < 0) continue;
    double v = 0.0;
    const int64_t a = ia[e], b = ib[e];
    if (a >= 0) v += alpha * av[a];
    if (b >= 0) v += beta * bv[b];
    L[d] = (T)v;
    local_max = fmax(local_max, fabs(v));
  }
  // wave reduction then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) local_max = fmax(local_max, __shfl_xor(local_max, off));
  if ((threadIdx.x & 63) == 0)
    atomicMax(&cnt[kCntKnorm], (unsigned long long)__double_as_longlong(local_max));
}

// ------------------------------------------------------------------------ Bunch-Kaufman in LDS

struct PivotCounts {
  int neg, zero, pos, two, exact_zero;
};

// Factorizes the leading ncand columns of the symmetric matrix F (LDS, column-major, leading
// dimension ld, lower triangle referenced, order mrows) by Bunch-Kaufman pivoting with candidate
// rows and columns restricted to [k, ncand). The Schur update covers rows/columns below up to
// mrows. On exit F holds L (unit, strictly lower; the (k+1, k) entry of a 2x2 pivot holds D's
// off-diagonal), D on the diagonal; perm[i] = original index of pivot position i; dv / dof / pt
// the pivot data. Called by all threads of the workgroup.
template <class T>
__device__ void bk_factor_lds(T* F, int ld, int ncand, int mrows, int* perm, T* dv, T* dof, int8_t* pt, T* t1,
                              T* t2, double tau, bool detect, PivotCounts& pc, const double* below) {
  __shared__ int s_kind, s_kp;
  __shared__ T s_inv[3];
  __shared__ double s_colmax[2];
  const int tid = threadIdx.x;
  const int nth = blockDim.x;
  for (int i = tid; i < ncand; i += nth) perm[i] = i;
  __syncthreads();
  int k = 0;
  while (k < ncand) {
    if (tid < 64) {
      const int lane = tid;
      const double absakk = fabs((double)F[k + k * ld]);
      double colmax = 0.0;
      int imax = -1;
      for (int i = k + 1 + lane; i < ncand; i += 64) {
        const double v = fabs((double)F[i + k * ld]);
        if (v > colmax) {
          colmax = v;
          imax = i;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(colmax, off);
        const int oi = __shfl_xor(imax, off);
        if (ov > colmax || (ov == colmax && oi >= 0 && (imax < 0 || oi < imax))) {
          colmax = ov;
          imax = oi;
        }
      }
      int kind = 1, kp = k;
      if (!(absakk >= kBkAlpha * colmax) && imax > k) {
        double rowmax = 0.0;
        for (int j = k + lane; j < ncand; j += 64) {
          if (j == imax) continue;
          const double v = j < imax ? fabs((double)F[imax + j * ld]) : fabs((double)F[j + imax * ld]);
          rowmax = fmax(rowmax, v);
        }
        for (int off = 32; off > 0; off >>= 1) rowmax = fmax(rowmax, __shfl_xor(rowmax, off));
        if (absakk * rowmax >= kBkAlpha * colmax * colmax) {
          kind = 1;
          kp = k;
        } else if (fabs((double)F[imax + imax * ld]) >= kBkAlpha * rowmax) {
          kind = 1;
          kp = imax;
        } else {
          kind = 2;
          kp = imax;
        }
      }
      if (lane == 0) {
        s_kind = kind;
        s_kp = kp;
      }
    }
    __syncthreads();
    const int kind = s_kind;
    const int kp = s_kp;
    const int kk = kind == 1 ? k : k + 1;
    if (kp != kk) {
      // symmetric interchange of positions p = kk < q = kp over the lower triangle of order mrows
      const int p = kk, q = kp;
      const int n_left = p;               // rows p, q of columns [0, p)
      const int n_mid = q - p - 1;        // A(j, p) <-> A(q, j), p < j < q
      const int n_below = mrows - q - 1;  // A(i, p) <-> A(i, q), i > q
      const int total = n_left + n_mid + n_below + 1;
      for (int t = tid; t < total; t += nth) {
        if (t < n_left) {
          const int j = t;
          const T a = F[p + j * ld];
          F[p + j * ld] = F[q + j * ld];
          F[q + j * ld] = a;
        } else if (t < n_left + n_mid) {
          const int j = p + 1 + (t - n_left);
          const T a = F[j + p * ld];
          F[j + p * ld] = F[q + j * ld];
          F[q + j * ld] = a;
        } else if (t < n_left + n_mid + n_below) {
          const int i = q + 1 + (t - n_left - n_mid);
          const T a = F[i + p * ld];
          F[i + p * ld] = F[i + q * ld];
          F[i + q * ld] = a;
        } else {
          const T a = F[p + p * ld];
          F[p + p * ld] = F[q + q * ld];
          F[q + q * ld] = a;
          const int pp = perm[p];
          perm[p] = perm[q];
          perm[q] = pp;
        }
      }
      __syncthreads();
    }
    // Null-pivot test on the whole remaining column (rows of the front below the pivot block,
    // contribution rows included, plus the rows below this LDS block for the blocked path): the
    // reference's rule flags a pivot whose remaining row is below tau (CNTL(3)), so a tiny pivot
    // of a nearly singular fully-summed block that still couples to its contribution rows is a
    // valid (if large-growth) pivot, not a singularity.
    if (tid < 64) {
      const int s0 = kind == 1 ? k + 1 : k + 2;
      double m0 = 0.0, m1 = 0.0;
      for (int i = s0 + tid; i < mrows; i += 64) {
        m0 = fmax(m0, fabs((double)F[i + k * ld]));
        if (kind == 2) m1 = fmax(m1, fabs((double)F[i + (k + 1) * ld]));
      }
      for (int off = 32; off > 0; off >>= 1) {
        m0 = fmax(m0, __shfl_xor(m0, off));
        m1 = fmax(m1, __shfl_xor(m1, off));
      }
      if (tid == 0) {
        if (below) {
          m0 = fmax(m0, below[perm[k]]);
          if (kind == 2) m1 = fmax(m1, below[perm[k + 1]]);
        }
        s_colmax[0] = m0;
        s_colmax[1] = m1;
      }
    }
    if (kind == 1) {
      if (tid == 0) {
        const double d = (double)F[k + k * ld];
        const bool coupled = s_colmax[0] > tau;
        bool null_piv = d == 0.0;
        if (d == 0.0) ++pc.exact_zero;
        if (detect && fabs(d) <= tau && !coupled) null_piv = true;
        if (null_piv) {
          ++pc.zero;
          pt[k] = kPivNull;
          s_inv[0] = (T)0;
        } else {
          if (d < 0) ++pc.neg;
          else ++pc.pos;
          pt[k] = kPiv1;
          s_inv[0] = (T)(1.0 / d);
        }
        dv[k] = (T)d;
        dof[k] = (T)0;
      }
      for (int i = k + 1 + tid; i < mrows; i += nth) t1[i] = F[i + k * ld];
      __syncthreads();
      const T inv = s_inv[0];
      if (inv != (T)0) {
        // trailing update of the lower triangle rows/cols (k, mrows)
        const int nt = mrows - k - 1;
        const int cg = nth / 64;  // column groups
        const int lane = tid & 63, grp = tid / 64;
        for (int jj = grp; jj < nt; jj += cg) {
          const int j = k + 1 + jj;
          const T lj = t1[j] * inv;
          for (int i = j + lane; i < mrows; i += 64) F[i + j * ld] -= t1[i] * lj;
        }
      }
      for (int i = k + 1 + tid; i < mrows; i += nth) F[i + k * ld] = t1[i] * inv;
      __syncthreads();
      k += 1;
    } else {
      if (tid == 0) {
        const double a = (double)F[k + k * ld], b = (double)F[k + 1 + k * ld], c = (double)F[k + 1 + (k + 1) * ld];
        const double det = a * c - b * b;
        const double mean = 0.5 * (a + c), rad = sqrt(0.25 * (a - c) * (a - c) + b * b);
        const double e1 = mean - rad, e2 = mean + rad;  // eigenvalues, e1 <= e2
        int nz = 0;
        if (det == 0.0) ++pc.exact_zero;
        const bool coupled = fmax(s_colmax[0], s_colmax[1]) > tau;
        for (double e : {e1, e2}) {
          const bool z = (detect && fabs(e) <= tau && !coupled) || e == 0.0;
          if (z) ++nz;
          else if (e < 0) ++pc.neg;
          else ++pc.pos;
        }
        pc.zero += nz;
        ++pc.two;
        if (nz > 0 || det == 0.0) {
          pt[k] = kPivNull;
          pt[k + 1] = kPivNull;
          s_inv[0] = s_inv[1] = s_inv[2] = (T)0;
        } else {
          pt[k] = kPiv2First;
          pt[k + 1] = kPiv2Second;
          s_inv[0] = (T)(c / det);
          s_inv[1] = (T)(-b / det);
          s_inv[2] = (T)(a / det);
        }
        dv[k] = (T)a;
        dv[k + 1] = (T)c;
        dof[k] = (T)b;
        dof[k + 1] = (T)0;
      }
      for (int i = k + 2 + tid; i < mrows; i += nth) {
        t1[i] = F[i + k * ld];
        t2[i] = F[i + (k + 1) * ld];
      }
      __syncthreads();
      const T i11 = s_inv[0], i12 = s_inv[1], i22 = s_inv[2];
      const bool live = !(i11 == (T)0 && i12 == (T)0 && i22 == (T)0);
      if (live) {
        const int nt = mrows - k - 2;
        const int cg = nth / 64;
        const int lane = tid & 63, grp = tid / 64;
        for (int jj = grp; jj < nt; jj += cg) {
          const int j = k + 2 + jj;
          const T l1j = t1[j] * i11 + t2[j] * i12;
          const T l2j = t1[j] * i12 + t2[j] * i22;
          for (int i = j + lane; i < mrows; i += 64) F[i + j * ld] -= t1[i] * l1j + t2[i] * l2j;
        }
      }
      for (int i = k + 2 + tid; i < mrows; i += nth) {
        F[i + k * ld] = t1[i] * i11 + t2[i] * i12;
        F[i + (k + 1) * ld] = t1[i] * i12 + t2[i] * i22;
      }
      __syncthreads();
      k += 2;
    }
  }
}

template <class T>
__device__ void flush_counts(const PivotCounts& pc, unsigned long long* cnt) {
  if (threadIdx.x == 0) {
    if (pc.neg) atomicAdd(&cnt[kCntNeg], (unsigned long long)pc.neg);
    if (pc.zero) atomicAdd(&cnt[kCntZero], (unsigned long long)pc.zero);
    if (pc.pos) atomicAdd(&cnt[kCntPos], (unsigned long long)pc.pos);
    if (pc.two) atomicAdd(&cnt[kCnt2x2], (unsigned long long)pc.two);
    if (pc.exact_zero) atomicAdd(&cnt[kCntExactZero], (unsigned long long)pc.exact_zero);
  }
}

// ------------------------------------------------------------------------------ small fronts

// One workgroup per front of the list; the whole front (m x m) in dynamic LDS.
template <class T>
__global__ void __launch_bounds__(kThreads) k_factor_small(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                          double null_tol, int detect) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int s = list[blockIdx.x];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int cb = m - ncol;
  const int64_t poff = S.panel_off[s];
  T* F = reinterpret_cast<T*>(smem_raw);
  T* t1 = F + (size_t)m * m;
  T* t2 = t1 + m;
  T* dv = t2 + m;
  T* dof = dv + m;
  int* perm = reinterpret_cast<int*>(dof + m);
  int8_t* pt = reinterpret_cast<int8_t*>(perm + m);
  const int tid = threadIdx.x;
  for (int idx = tid; idx < m * m; idx += blockDim.x) {
    const int c = idx / m;
    F[idx] = c < ncol ? N.L[poff + idx] : (T)0;
  }
  __syncthreads();
  for (int q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q) {
    const int ch = S.child_list[q];
    const int cbc = S.sn_m[ch] - S.sn_ncol[ch];
    const int32_t* rel = S.relmap + S.sn_rowoff[ch] + S.sn_ncol[ch];
    const T* cbp = N.CB + S.cb_off[ch];
    for (int idx = tid; idx < cbc * cbc; idx += blockDim.x) {
      const int i = idx % cbc, j = idx / cbc;
      if (i >= j) F[rel[i] + rel[j] * m] += cbp[idx];
    }
    __syncthreads();
  }
  const double tau = null_tol * as_double_bits(N.cnt[kCntKnorm]);
  PivotCounts pc{0, 0, 0, 0, 0};
  bk_factor_lds<T>(F, m, ncol, m, perm, dv, dof, pt, t1, t2, tau, detect != 0, pc, nullptr);
  for (int idx = tid; idx < m * ncol; idx += blockDim.x) N.L[poff + idx] = F[idx];
  T* cbs = N.CB + S.cb_off[s];
  for (int idx = tid; idx < cb * cb; idx += blockDim.x) {
    const int i = idx % cb, j = idx / cb;
    cbs[idx] = i >= j ? F[(ncol + i) + (ncol + j) * m] : (T)0;
  }
  for (int i = tid; i < ncol; i += blockDim.x) {
    N.dval[first + i] = dv[i];
    N.doff[first + i] = dof[i];
    N.ptype[first + i] = pt[i];
    N.piv[first + i] = perm[i];
    N.ipiv[first + perm[i]] = i;
  }
  flush_counts<T>(pc, N.cnt);
}

// ------------------------------------------------------------------------------ large fronts

constexpr int kEaCols = 16;  // parent columns per extend-add workgroup

template <class T>
__global__ void __launch_bounds__(kThreads) k_extend_add_large(DevSym S, DevNum<T> N, const int32_t* __restrict__ list) {
  const int s = list[blockIdx.y];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s];
  const int cb = m - ncol;
  const int a = blockIdx.x * kEaCols;
  if (a >= m) return;
  const int b = min(m, a + kEaCols);
  T* P = N.L + S.panel_off[s];
  T* C = N.CB + S.cb_off[s];
  for (int q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q) {
    const int ch = S.child_list[q];
    const int cbc = S.sn_m[ch] - S.sn_ncol[ch];
    const int32_t* rel = S.relmap + S.sn_rowoff[ch] + S.sn_ncol[ch];
    const T* cbp = N.CB + S.cb_off[ch];
    // child columns mapping into [a, b): contiguous because rel is increasing
    int lo = 0, hi = cbc;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (rel[mid] < a) lo = mid + 1;
      else hi = mid;
    }
    const int j0 = lo;
    hi = cbc;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (rel[mid] < b) lo = mid + 1;
      else hi = mid;
    }
    const int j1 = lo;
    for (int j = j0; j < j1; ++j) {
      const int pc = rel[j];
      const T* col = cbp + (int64_t)j * cbc;
      if (pc < ncol) {
        T* dst = P + (int64_t)pc * m;
        for (int i = j + threadIdx.x; i < cbc; i += blockDim.x) dst[rel[i]] += col[i];
      } else {
        T* dst = C + (int64_t)(pc - ncol) * cb - ncol;
        for (int i = j + threadIdx.x; i < cbc; i += blockDim.x) dst[rel[i]] += col[i];
      }
    }
    __syncthreads();
  }
}

// Bunch-Kaufman on the nb x nb diagonal block of panel p of each front (one workgroup per front).
template <class T>
__global__ void __launch_bounds__(kThreads) k_bk_diag(DevSym S, DevNum<T> N, const int32_t* __restrict__ list, int p,
                                                     int nb, int32_t* __restrict__ bperm, double null_tol,
                                                     int detect) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int f = blockIdx.x;
  const int s = list[f];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int c0 = p * nb;
  if (c0 >= ncol) return;
  const int w = min(nb, ncol - c0);
  T* P = N.L + S.panel_off[s];
  T* F = reinterpret_cast<T*>(smem_raw);
  T* t1 = F + nb * nb;
  T* t2 = t1 + nb;
  T* dv = t2 + nb;
  T* dof = dv + nb;
  int* perm = reinterpret_cast<int*>(dof + nb);
  int8_t* pt = reinterpret_cast<int8_t*>(perm + nb);
  const int tid = threadIdx.x;
  __shared__ double below[64];
  for (int idx = tid; idx < w * w; idx += blockDim.x) {
    const int i = idx % w, j = idx / w;
    F[i + j * w] = i >= j ? P[(int64_t)(c0 + j) * m + c0 + i] : (T)0;
  }
  // column maxima of the panel rows below the block (the null-pivot test's remaining rows)
  {
    const int lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    for (int j = wave; j < w; j += nw) {
      double mx = 0.0;
      for (int r = c0 + w + lane; r < m; r += 64) mx = fmax(mx, fabs((double)P[(int64_t)(c0 + j) * m + r]));
      for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
      if (lane == 0) below[j] = mx;
    }
  }
  __syncthreads();
  const double tau = null_tol * as_double_bits(N.cnt[kCntKnorm]);
  PivotCounts pc{0, 0, 0, 0, 0};
  bk_factor_lds<T>(F, w, w, w, perm, dv, dof, pt, t1, t2, tau, detect != 0, pc, below);
  for (int idx = tid; idx < w * w; idx += blockDim.x) {
    const int i = idx % w, j = idx / w;
    if (i >= j) P[(int64_t)(c0 + j) * m + c0 + i] = F[i + j * w];
  }
  for (int i = tid; i < w; i += blockDim.x) {
    N.dval[first + c0 + i] = dv[i];
    N.doff[first + c0 + i] = dof[i];
    N.ptype[first + c0 + i] = pt[i];
    N.piv[first + c0 + i] = c0 + perm[i];
    N.ipiv[first + c0 + perm[i]] = c0 + i;
    bperm[(int64_t)f * nb + i] = perm[i];
  }
  (void)m;
  flush_counts<T>(pc, N.cnt);
}

constexpr int kPanelRows = 64;  // rows per panel-solve workgroup

// Panel solve below the diagonal block: for rows r in [c0 + w, m) of the panel columns
// [c0, c0 + w): apply the block's column permutation, W = A21 L11^-T (stored in the W workspace),
// L21 = W D^-1 (stored in the panel). Workgroups beyond the row chunks permute rows
// [c0, c0 + w) of the earlier panel columns [0, c0).
template <class T>
__global__ void __launch_bounds__(kThreads) k_panel_apply(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                         int p, int nb, const int32_t* __restrict__ bperm) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int f = blockIdx.y;
  const int s = list[f];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int c0 = p * nb;
  if (c0 >= ncol) return;
  const int w = min(nb, ncol - c0);
  const int c1 = c0 + w;
  T* P = N.L + S.panel_off[s];
  const int32_t* bp = bperm + (int64_t)f * nb;
  const int nrows = m - c1;
  const int nrc = (nrows + kPanelRows - 1) / kPanelRows;
  const int tid = threadIdx.x;
  if ((int)blockIdx.x >= nrc) {
    // row interchanges of the earlier columns: columns handled in groups of kThreads / w
    const int cpg = max(1, kThreads / w);
    const int j0 = (blockIdx.x - nrc) * cpg;
    if (j0 >= c0) return;
    T* tile = reinterpret_cast<T*>(smem_raw);  // w x cpg
    const int i = tid % w, jj = tid / w;
    const int j = j0 + jj;
    const bool act = jj < cpg && j < c0 && i < w && tid < w * cpg;
    if (act) tile[i + jj * w] = P[(int64_t)j * m + c0 + bp[i]];
    __syncthreads();
    if (act) P[(int64_t)j * m + c0 + i] = tile[i + jj * w];
    return;
  }
  const int r0 = c1 + blockIdx.x * kPanelRows;
  const int rn = min(kPanelRows, m - r0);
  T* L11 = reinterpret_cast<T*>(smem_raw);  // w x w
  T* X = L11 + nb * nb;                      // rn x w, ld kPanelRows
  __shared__ T dinv[3 * 64];
  __shared__ int8_t ptl[64];
  for (int idx = tid; idx < w * w; idx += blockDim.x) {
    const int i = idx % w, j = idx / w;
    T v = i > j ? P[(int64_t)(c0 + j) * m + c0 + i] : (T)0;
    L11[idx] = v;
  }
  for (int i = tid; i < w; i += blockDim.x) ptl[i] = N.ptype[first + c0 + i];
  __syncthreads();
  // the (j+1, j) entry of a 2x2 pivot is D's off-diagonal, not part of L11
  for (int j = tid; j < w; j += blockDim.x)
    if ((ptl[j] == kPiv2First || ptl[j] == kPivNull) && j + 1 < w) L11[(j + 1) + j * w] = (T)0;
  if (tid < w) {
    const int j = tid;
    const int8_t t = ptl[j];
    const T d = N.dval[first + c0 + j];
    if (t == kPiv1) {
      dinv[3 * j] = (T)1 / d;
      dinv[3 * j + 1] = dinv[3 * j + 2] = (T)0;
    } else if (t == kPiv2First) {
      const double a = (double)d, c = (double)N.dval[first + c0 + j + 1], b = (double)N.doff[first + c0 + j];
      const double det = a * c - b * b;
      dinv[3 * j] = (T)(c / det);
      dinv[3 * j + 1] = (T)(-b / det);
      dinv[3 * j + 2] = (T)(a / det);
    } else {
      dinv[3 * j] = dinv[3 * j + 1] = dinv[3 * j + 2] = (T)0;
    }
  }
  // gather the permuted columns of A21
  for (int idx = tid; idx < rn * w; idx += blockDim.x) {
    const int r = idx % rn, j = idx / rn;
    X[r + j * kPanelRows] = P[(int64_t)(c0 + bp[j]) * m + r0 + r];
  }
  __syncthreads();
  // W = A21 L11^-T, right-looking over the w columns
  for (int j = 0; j < w; ++j) {
    const int nt = w - j - 1;
    for (int idx = tid; idx < rn * nt; idx += blockDim.x) {
      const int r = idx % rn, t = j + 1 + idx / rn;
      X[r + t * kPanelRows] -= X[r + j * kPanelRows] * L11[t + j * w];
    }
    __syncthreads();
  }
  T* Wp = N.W + S.w_off[s];
  for (int idx = tid; idx < rn * w; idx += blockDim.x) {
    const int r = idx % rn, j = idx / rn;
    Wp[(int64_t)(c0 + j) * m + r0 + r] = X[r + j * kPanelRows];
  }
  // L21 = W D^-1
  for (int idx = tid; idx < rn * w; idx += blockDim.x) {
    const int r = idx % rn, j = idx / rn;
    const int8_t t = ptl[j];
    T v;
    if (t == kPiv1) {
      v = X[r + j * kPanelRows] * dinv[3 * j];
    } else if (t == kPiv2First) {
      v = X[r + j * kPanelRows] * dinv[3 * j] + X[r + (j + 1) * kPanelRows] * dinv[3 * j + 1];
    } else if (t == kPiv2Second) {
      v = X[r + (j - 1) * kPanelRows] * dinv[3 * (j - 1) + 1] + X[r + j * kPanelRows] * dinv[3 * (j - 1) + 2];
    } else {
      v = (T)0;
    }
    P[(int64_t)(c0 + j) * m + r0 + r] = v;
  }
}

// C -= A B^T restricted to the lower triangle, per front:
//   mode 0 (trailing update of panel p): rows [c1, m) x cols [c1, ncol) of the panel,
//          A = W(:, c0:c1), B = L(:, c0:c1);
//   mode 1 (contribution block): rows/cols [ncol, m) into CB, A = W(:, 0:ncol), B = L(:, 0:ncol).
// 64 x 64 output tiles, 256 threads with 4 x 4 outputs each, K staged through LDS in steps of 16.
constexpr int kTile = 64;
constexpr int kKStep = 16;

template <class T>
__global__ void __launch_bounds__(kThreads) k_gemm_update(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                         int mode, int p, int nb) {
  const int s = list[blockIdx.y];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s];
  int row0, col0, k0, kn;
  if (mode == 0) {
    const int c0 = p * nb;
    if (c0 >= ncol) return;
    const int c1 = min(ncol, c0 + nb);
    if (c1 >= ncol) return;
    row0 = c1;
    col0 = c1;
    k0 = c0;
    kn = c1 - c0;
  } else {
    if (m == ncol) return;
    row0 = ncol;
    col0 = ncol;
    k0 = 0;
    kn = ncol;
  }
  const int nr = m - row0;
  const int ncl = mode == 0 ? ncol - col0 : m - col0;
  const int tr = (nr + kTile - 1) / kTile, tc = (ncl + kTile - 1) / kTile;
  const int tile = blockIdx.x;
  if (tile >= tr * tc) return;
  const int ti = tile / tc, tj = tile % tc;
  if (tj * kTile > ti * kTile + kTile - 1 + (row0 - col0)) return;  // entirely above the diagonal
  const T* A = N.W + S.w_off[s];
  const T* B = N.L + S.panel_off[s];
  __shared__ T As[kKStep][kTile + 1];
  __shared__ T Bs[kKStep][kTile + 1];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (T)0;
  const int rbase = row0 + ti * kTile, cbase = col0 + tj * kTile;
  for (int kk = 0; kk < kn; kk += kKStep) {
    for (int idx = tid; idx < kKStep * kTile; idx += kThreads) {
      const int r = idx % kTile, t = idx / kTile;
      const int gr = rbase + r, gc = cbase + r, gk = k0 + kk + t;
      const bool kin = kk + t < kn;
      As[t][r] = (kin && gr < m) ? A[(int64_t)gk * m + gr] : (T)0;
      Bs[t][r] = (kin && gc < (mode == 0 ? ncol : m)) ? B[(int64_t)gk * m + gc] : (T)0;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kKStep; ++t) {
      T av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[t][tx + 16 * a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[t][ty + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
    }
    __syncthreads();
  }
  const int cb = m - ncol;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int gr = rbase + tx + 16 * a, gc = cbase + ty + 16 * b;
      if (gr >= m || gr < gc) continue;
      if (mode == 0) {
        if (gc >= ncol) continue;
        N.L[S.panel_off[s] + (int64_t)gc * m + gr] -= acc[a][b];
      } else {
        N.CB[S.cb_off[s] + (int64_t)(gc - ncol) * cb + (gr - ncol)] -= acc[a][b];
      }
    }
}

// ------------------------------------------------------------------------------------ solves
//
// Multifrontal forward / backward substitution with p right-hand sides (the reference's MUMPS
// JOB = 3, src/factor/mumps_factorization.cpp:458-527). Forward (levels bottom-up): a front
// gathers its own right-hand-side rows (pivot order), adds its children's update vectors, solves
// with the unit lower L11 and passes U = v_cb - L21 y up; the pivot-order intermediate lives in Z
// (n x p). D^-1 is applied per pivot block. Backward (levels top-down): a front subtracts
// L21^T x_cb (ancestor rows, already final), solves with L11^T and scatters its rows into X.
// R is the right-hand-side type (double), T the factor type.

// L(i, j) of the pivot block, 0 at the (j+1, j) position of a 2x2 or null pivot pair.
template <class T>
__device__ __forceinline__ T l11_at(const T* P, int m, int i, int j, const int8_t* pt) {
  if (i == j + 1 && (pt[j] == kPiv2First || pt[j] == kPivNull)) return (T)0;
  return P[(int64_t)j * m + i];
}

template <class T, class R>
__global__ void __launch_bounds__(kThreads) k_solve_forward(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                           const R* __restrict__ X, int64_t ldx, int p, T* Z,
                                                           T* U, int64_t ldu) {
  const int s = list[blockIdx.x];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int cb = m - ncol;
  const int n = S.n;
  const int64_t roff = S.sn_rowoff[s];
  const int32_t* rows = S.rows_orig + roff;
  const T* P = N.L + S.panel_off[s];
  const int8_t* pt = N.ptype + first;
  T* zs = Z + first;
  T* us = U + roff + ncol;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < ncol * p; idx += blockDim.x) {
    const int i = idx % ncol, c = idx / ncol;
    zs[i + (int64_t)c * n] = (T)X[rows[N.piv[first + i]] + (int64_t)c * ldx];
  }
  for (int idx = tid; idx < cb * p; idx += blockDim.x) {
    const int i = idx % cb, c = idx / cb;
    us[i + (int64_t)c * ldu] = (T)0;
  }
  __syncthreads();
  for (int q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q) {
    const int ch = S.child_list[q];
    const int cbc = S.sn_m[ch] - S.sn_ncol[ch];
    const int64_t coff = S.sn_rowoff[ch] + S.sn_ncol[ch];
    const int32_t* rel = S.relmap + coff;
    const T* uc = U + coff;
    for (int idx = tid; idx < cbc * p; idx += blockDim.x) {
      const int i = idx % cbc, c = idx / cbc;
      const int loc = rel[i];
      const T v = uc[i + (int64_t)c * ldu];
      if (loc < ncol) zs[N.ipiv[first + loc] + (int64_t)c * n] += v;
      else us[loc - ncol + (int64_t)c * ldu] += v;
    }
    __syncthreads();
  }
  // L11 y = v, column by column
  for (int j = 0; j < ncol - 1; ++j) {
    const int nt = ncol - j - 1;
    for (int idx = tid; idx < nt * p; idx += blockDim.x) {
      const int i = j + 1 + idx % nt, c = idx / nt;
      const T l = l11_at(P, m, i, j, pt);
      if (l != (T)0) zs[i + (int64_t)c * n] -= l * zs[j + (int64_t)c * n];
    }
    __syncthreads();
  }
  // U = v_cb - L21 y
  for (int idx = tid; idx < cb * p; idx += blockDim.x) {
    const int i = idx % cb, c = idx / cb;
    T acc = (T)0;
    for (int j = 0; j < ncol; ++j) acc += P[(int64_t)j * m + ncol + i] * zs[j + (int64_t)c * n];
    us[i + (int64_t)c * ldu] -= acc;
  }
}

template <class T>
__global__ void k_solve_diag(int n, const T* __restrict__ dval, const T* __restrict__ doff,
                             const int8_t* __restrict__ ptype, T* Z, int p) {
  const int64_t total = (int64_t)n * p;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n);
    const int64_t c = idx / n;
    const int8_t t = ptype[i];
    T* z = Z + c * n;
    if (t == kPiv1) {
      z[i] = z[i] / dval[i];
    } else if (t == kPiv2First) {
      const T a = dval[i], b = doff[i], d = dval[i + 1];
      const T det = a * d - b * b;
      const T z1 = z[i], z2 = z[i + 1];
      z[i] = (d * z1 - b * z2) / det;
      z[i + 1] = (a * z2 - b * z1) / det;
    } else if (t == kPivNull) {
      z[i] = (T)0;
    }
  }
}

template <class T, class R>
__global__ void __launch_bounds__(kThreads) k_solve_backward(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                            R* X, int64_t ldx, int p, T* Z) {
  __shared__ T red[kThreads / 64][16];
  const int s = list[blockIdx.x];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int cb = m - ncol;
  const int n = S.n;
  const int64_t roff = S.sn_rowoff[s];
  const int32_t* rows = S.rows_orig + roff;
  const T* P = N.L + S.panel_off[s];
  const int8_t* pt = N.ptype + first;
  T* zs = Z + first;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  // y_j -= sum_i L21(i, j) x_cb(i): one wave per column
  if (cb > 0) {
    for (int j = wave; j < ncol; j += nw) {
      for (int c = 0; c < p; ++c) {
        T acc = (T)0;
        for (int i = lane; i < cb; i += 64)
          acc += P[(int64_t)j * m + ncol + i] * (T)X[rows[ncol + i] + (int64_t)c * ldx];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) zs[j + (int64_t)c * n] -= acc;
      }
    }
    __syncthreads();
  }
  // L11^T x = y, from the last pivot up: x_j = y_j - sum_{i > j} L(i, j) x_i
  for (int j = ncol - 2; j >= 0; --j) {
    const int nt = ncol - j - 1;
    for (int c0 = 0; c0 < p; c0 += 16) {
      const int pc = min(16, p - c0);
      T acc[16];
      for (int c = 0; c < pc; ++c) acc[c] = (T)0;
      for (int t = tid; t < nt; t += blockDim.x) {
        const int i = j + 1 + t;
        const T l = l11_at(P, m, i, j, pt);
        for (int c = 0; c < pc; ++c) acc[c] += l * zs[i + (int64_t)(c0 + c) * n];
      }
      for (int c = 0; c < pc; ++c) {
        T v = acc[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) red[wave][c] = v;
      }
      __syncthreads();
      if (tid < pc) {
        T v = (T)0;
        for (int w = 0; w < nw; ++w) v += red[w][tid];
        zs[j + (int64_t)(c0 + tid) * n] -= v;
      }
      __syncthreads();
    }
  }
  for (int idx = tid; idx < ncol * p; idx += blockDim.x) {
    const int i = idx % ncol, c = idx / ncol;
    X[rows[N.piv[first + i]] + (int64_t)c * ldx] = (R)zs[i + (int64_t)c * n];
  }
}

// ------------------------------------------------------------------------------ host engine

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

template <class V>
V* upload(const std::vector<V>& h) {
  V* d = nullptr;
  const size_t bytes = std::max<size_t>(1, h.size()) * sizeof(V);
  hip_check(hipMalloc(&d, bytes), "hipMalloc (analysis)");
  if (!h.empty()) hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice), "hipMemcpy (analysis)");
  return d;
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

int blocks_for(int64_t n, int per = 256, int cap = 65536) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(cap, (n + per - 1) / per));
}

}  // namespace

Engine::Engine(Symbolic sym, std::vector<int64_t> asm_dst, std::vector<int64_t> ia, std::vector<int64_t> ib,
               const EngineOptions& opt)
    : sym_(std::move(sym)), opt_(opt), h_asm_(std::move(asm_dst)), h_ia_(std::move(ia)), h_ib_(std::move(ib)) {
  if (opt_.panel_width < 8 || opt_.panel_width > 64) throw std::runtime_error("panel_width must be in [8, 64]");
  if (opt_.small_front_max < 1 || opt_.small_front_max > 128) throw std::runtime_error("small_front_max must be in [1, 128]");
  nnz_ = (int64_t)h_asm_.size();
  const int nsn = sym_.nsn;
  // W workspace offsets for large fronts
  h_w_off_.assign(nsn, -1);
  for (int s = 0; s < nsn; ++s) {
    if (sym_.sn_m[s] > opt_.small_front_max) {
      h_w_off_[s] = w_total_;
      w_total_ += (int64_t)sym_.sn_m[s] * sym_.sn_ncol[s];
    }
  }
  panel_total_ = sym_.panel_off[nsn];
  cb_total_ = sym_.cb_off[nsn];
  rows_total_ = (int64_t)sym_.rows.size();
  levels_.resize(sym_.nlevels);
  for (int l = 0; l < sym_.nlevels; ++l) {
    Level& L = levels_[l];
    for (int q = sym_.level_ptr[l]; q < sym_.level_ptr[l + 1]; ++q) {
      const int s = sym_.level_sn[q];
      L.h_all.push_back(s);
      if (sym_.sn_m[s] <= opt_.small_front_max) {
        L.h_small.push_back(s);
        L.small_maxm = std::max(L.small_maxm, sym_.sn_m[s]);
      } else {
        L.h_large.push_back(s);
        L.large_maxm = std::max(L.large_maxm, sym_.sn_m[s]);
        L.large_max_ncol = std::max(L.large_max_ncol, sym_.sn_ncol[s]);
        L.large_max_cb = std::max(L.large_max_cb, sym_.sn_m[s] - sym_.sn_ncol[s]);
      }
    }
    L.n_all = (int)L.h_all.size();
    L.n_small = (int)L.h_small.size();
    L.n_large = (int)L.h_large.size();
    n_small_ += L.n_small;
    n_large_ += L.n_large;
    max_large_per_level_ = std::max(max_large_per_level_, L.n_large);
  }
}

void Engine::ensure_device() {
  if (on_device_) return;
  std::vector<int32_t> rows_orig(sym_.rows.size());
  for (size_t k = 0; k < sym_.rows.size(); ++k) rows_orig[k] = sym_.order[sym_.rows[k]];
  d_first_ = upload(sym_.sn_first);
  d_ncol_ = upload(sym_.sn_ncol);
  d_m_ = upload(sym_.sn_m);
  d_child_ptr_ = upload(sym_.child_ptr);
  d_child_list_ = upload(sym_.child_list);
  d_relmap_ = upload(sym_.relmap);
  d_rows_orig_ = upload(rows_orig);
  d_rowoff_ = upload(sym_.sn_rowoff);
  d_panel_off_ = upload(sym_.panel_off);
  d_cb_off_ = upload(sym_.cb_off);
  d_w_off_ = upload(h_w_off_);
  d_asm_ = upload(h_asm_);
  d_ia_ = upload(h_ia_);
  d_ib_ = upload(h_ib_);
  for (Level& L : levels_) {
    L.d_all = upload(L.h_all);
    if (L.n_small) L.d_small = upload(L.h_small);
    if (L.n_large) L.d_large = upload(L.h_large);
  }
  on_device_ = true;
}

Engine::~Engine() {
  free_numeric();
  for (void* p : {(void*)d_first_, (void*)d_ncol_, (void*)d_m_, (void*)d_child_ptr_, (void*)d_child_list_,
                  (void*)d_relmap_, (void*)d_rows_orig_, (void*)d_rowoff_, (void*)d_panel_off_, (void*)d_cb_off_,
                  (void*)d_w_off_, (void*)d_asm_, (void*)d_ia_, (void*)d_ib_})
    dfree(p);
  for (auto& L : levels_) {
    dfree(L.d_all);
    dfree(L.d_small);
    dfree(L.d_large);
  }
}

void Engine::free_numeric() {
  for (void* p : {d_L_, d_CB_, d_W_, d_dval_, d_doff_, (void*)d_ptype_, (void*)d_piv_, (void*)d_ipiv_,
                  (void*)d_bperm_, (void*)d_cnt_, d_Z_, d_U_})
    dfree(p);
  d_L_ = d_CB_ = d_W_ = d_dval_ = d_doff_ = d_Z_ = d_U_ = nullptr;
  d_ptype_ = nullptr;
  d_piv_ = d_ipiv_ = d_bperm_ = nullptr;
  d_cnt_ = nullptr;
  ws_cols_ = 0;
  precision_ = -1;
  valid_ = false;
}

void Engine::release() { free_numeric(); }

void Engine::alloc_numeric(int precision) {
  if (precision_ == precision && d_L_) return;
  free_numeric();
  const size_t ts = precision == 0 ? sizeof(double) : sizeof(float);
  const int64_t n = sym_.n;
  hip_check(hipMalloc(&d_L_, std::max<int64_t>(1, panel_total_) * ts), "hipMalloc (factor panels)");
  hip_check(hipMalloc(&d_CB_, std::max<int64_t>(1, cb_total_) * ts), "hipMalloc (contribution blocks)");
  hip_check(hipMalloc(&d_W_, std::max<int64_t>(1, w_total_) * ts), "hipMalloc (panel workspace)");
  hip_check(hipMalloc(&d_dval_, n * ts), "hipMalloc (D)");
  hip_check(hipMalloc(&d_doff_, n * ts), "hipMalloc (D)");
  hip_check(hipMalloc(&d_ptype_, n), "hipMalloc (pivots)");
  hip_check(hipMalloc(&d_piv_, n * sizeof(int32_t)), "hipMalloc (pivots)");
  hip_check(hipMalloc(&d_ipiv_, n * sizeof(int32_t)), "hipMalloc (pivots)");
  hip_check(hipMalloc(&d_bperm_, std::max(1, max_large_per_level_) * (size_t)opt_.panel_width * sizeof(int32_t)),
            "hipMalloc (pivots)");
  hip_check(hipMalloc(&d_cnt_, kCntSlots * sizeof(unsigned long long)), "hipMalloc (counters)");
  precision_ = precision;
}

template <class T>
InertiaResult Engine::factorize_t(double alpha, double beta, const double* av, const double* bv, double null_tol,
                                  hipStream_t st) {
  DevSym S{sym_.n, sym_.nsn, d_first_, d_ncol_, d_m_, d_rowoff_, d_panel_off_, d_cb_off_, d_w_off_,
           d_child_ptr_, d_child_list_, d_relmap_, d_rows_orig_};
  DevNum<T> N{(T*)d_L_, (T*)d_CB_, (T*)d_W_, (T*)d_dval_, (T*)d_doff_, d_ptype_, d_piv_, d_ipiv_, d_cnt_};
  const int nb = opt_.panel_width;
  const int detect = null_tol > 0 ? 1 : 0;
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks_for(panel_total_)), dim3(256), 0, st, N.L, panel_total_, (T)0);
  if (cb_total_ > 0)
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks_for(cb_total_)), dim3(256), 0, st, N.CB, cb_total_, (T)0);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, d_cnt_, (int)kCntSlots);
  if (nnz_ > 0)
    hipLaunchKernelGGL(k_assemble_original<T>, dim3(blocks_for(nnz_, 256, 8192)), dim3(256), 0, st, nnz_, d_asm_, d_ia_,
                       d_ib_, av, bv, alpha, beta, N.L, d_cnt_);
  for (const Level& L : levels_) {
    if (L.n_small) {
      const int m = L.small_maxm;
      const size_t shm = (size_t)m * m * sizeof(T) + 4 * (size_t)m * sizeof(T) + (size_t)m * sizeof(int) + m + 16;
      hipLaunchKernelGGL(k_factor_small<T>, dim3(L.n_small), dim3(kThreads), shm, st, S, N, L.d_small, null_tol,
                         detect);
    }
    if (L.n_large) {
      hipLaunchKernelGGL(k_extend_add_large<T>, dim3((L.large_maxm + kEaCols - 1) / kEaCols, L.n_large),
                         dim3(kThreads), 0, st, S, N, L.d_large);
      const int npan = (L.large_max_ncol + nb - 1) / nb;
      const size_t shm_bk = (size_t)nb * nb * sizeof(T) + 4 * (size_t)nb * sizeof(T) + nb * sizeof(int) + nb + 16;
      const size_t shm_pa = ((size_t)nb * nb + (size_t)kPanelRows * nb) * sizeof(T);
      const int nrc = (L.large_maxm + kPanelRows - 1) / kPanelRows;
      const int cpg = std::max(1, kThreads / nb);
      const int nsw = (L.large_max_ncol + cpg - 1) / cpg;
      for (int p = 0; p < npan; ++p) {
        hipLaunchKernelGGL(k_bk_diag<T>, dim3(L.n_large), dim3(kThreads), shm_bk, st, S, N, L.d_large, p, nb, d_bperm_,
                           null_tol, detect);
        hipLaunchKernelGGL(k_panel_apply<T>, dim3(nrc + nsw, L.n_large), dim3(kThreads), shm_pa, st, S, N, L.d_large, p,
                           nb, (const int32_t*)d_bperm_);
        const int rem = L.large_maxm - (p + 1) * nb;
        if (rem > 0 && (p + 1) * nb < L.large_max_ncol) {
          const int t = (rem + kTile - 1) / kTile;
          const int tc = (L.large_max_ncol - (p + 1) * nb + kTile - 1) / kTile;
          hipLaunchKernelGGL(k_gemm_update<T>, dim3(t * tc, L.n_large), dim3(kThreads), 0, st, S, N, L.d_large, 0, p,
                             nb);
        }
      }
      if (L.large_max_cb > 0) {
        const int t = (L.large_max_cb + kTile - 1) / kTile;
        hipLaunchKernelGGL(k_gemm_update<T>, dim3(t * t, L.n_large), dim3(kThreads), 0, st, S, N, L.d_large, 1, 0, nb);
      }
    }
  }
  hip_check(hipGetLastError(), "factorization kernels");
  unsigned long long h[kCntSlots];
  hip_check(hipMemcpyAsync(h, d_cnt_, sizeof(h), hipMemcpyDeviceToHost, st), "hipMemcpyAsync (counters)");
  hip_check(hipStreamSynchronize(st), "factorization");
  InertiaResult r;
  r.neg = (int64_t)h[kCntNeg];
  r.zero = (int64_t)h[kCntZero];
  r.pos = (int64_t)h[kCntPos];
  r.two = (int64_t)h[kCnt2x2];
  r.exact_zero = (int64_t)h[kCntExactZero];
  r.singular = r.zero > 0;
  if (r.neg + r.zero + r.pos != sym_.n) throw std::runtime_error("factorization: pivot counts do not add up to n");
  return r;
}

InertiaResult Engine::factorize(double alpha, double beta, const double* av, const double* bv, int precision,
                                double null_tol, hipStream_t stream) {
  valid_ = false;
  ensure_device();
  alloc_numeric(precision);
  InertiaResult r = precision == 0 ? factorize_t<double>(alpha, beta, av, bv, null_tol, stream)
                                   : factorize_t<float>(alpha, beta, av, bv, null_tol, stream);
  valid_ = !r.singular;
  return r;
}

template <class T>
void Engine::solve_t(double* x, int64_t nrhs, int64_t ldx, hipStream_t st) {
  DevSym S{sym_.n, sym_.nsn, d_first_, d_ncol_, d_m_, d_rowoff_, d_panel_off_, d_cb_off_, d_w_off_,
           d_child_ptr_, d_child_list_, d_relmap_, d_rows_orig_};
  DevNum<T> N{(T*)d_L_, (T*)d_CB_, (T*)d_W_, (T*)d_dval_, (T*)d_doff_, d_ptype_, d_piv_, d_ipiv_, d_cnt_};
  if (ws_cols_ < nrhs) {
    dfree(d_Z_);
    dfree(d_U_);
    d_Z_ = d_U_ = nullptr;
    hip_check(hipMalloc(&d_Z_, (size_t)sym_.n * nrhs * sizeof(T)), "hipMalloc (solve workspace)");
    hip_check(hipMalloc(&d_U_, (size_t)std::max<int64_t>(1, rows_total_) * nrhs * sizeof(T)),
              "hipMalloc (solve workspace)");
    ws_cols_ = nrhs;
  }
  T* Z = (T*)d_Z_;
  T* U = (T*)d_U_;
  const int p = (int)nrhs;
  for (const Level& L : levels_)
    hipLaunchKernelGGL((k_solve_forward<T, double>), dim3(L.n_all), dim3(kThreads), 0, st, S, N, L.d_all,
                       (const double*)x, ldx, p, Z, U, rows_total_);
  hipLaunchKernelGGL(k_solve_diag<T>, dim3(blocks_for((int64_t)sym_.n * p)), dim3(256), 0, st, sym_.n, N.dval,
                     N.doff, N.ptype, Z, p);
  for (int l = (int)levels_.size() - 1; l >= 0; --l) {
    const Level& L = levels_[l];
    hipLaunchKernelGGL((k_solve_backward<T, double>), dim3(L.n_all), dim3(kThreads), 0, st, S, N, L.d_all, x, ldx, p,
                       Z);
  }
  hip_check(hipGetLastError(), "solve kernels");
}

void Engine::solve(double* x, int64_t nrhs, int64_t ldx, hipStream_t stream) {
  if (!valid_) throw std::runtime_error("solve without a valid factorization");
  if (nrhs <= 0) return;
  if (ldx < sym_.n) throw std::runtime_error("solve: ldx < n");
  if (precision_ == 0) solve_t<double>(x, nrhs, ldx, stream);
  else solve_t<float>(x, nrhs, ldx, stream);
}


}  // namespace hsg
