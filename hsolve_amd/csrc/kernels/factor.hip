// Multifrontal LDL^T with Bunch-Kaufman pivoting inside fronts, on CDNA4 (gfx950).
//
// Replaces the numeric factorization the reference delegates to MUMPS (SYM = 2, JOB = 2;
// src/factor/mumps_factorization.cpp:299-356) and the inertia it reads from INFOG(12)/(28)
// (:358-381). Fronts are processed level by level of the assembly tree (children before parents):
//   * small fronts (order m <= small_max): one workgroup per front, the whole front in LDS:
//     load the panel (original entries were scattered there beforehand), extend-add the children's
//     contribution blocks, factor the pivot columns with Bunch-Kaufman 1x1/2x2 pivots restricted
//     to the front's fully-summed block, write L, D and the contribution block back;
//   * large fronts: extend-add of the children into the pivot columns, by column ranges (each
//     parent column owned by one workgroup, children in order: deterministic), then per panel of
//     nb pivot columns a one-workgroup Bunch-Kaufman factorization of the nb x nb diagonal block,
//     a multi-workgroup panel solve that forms W = A21 L11^-T and L21 = W D^-1, and a trailing
//     GEMM update of the remaining pivot columns; finally one GEMM stores the contribution block
//     CB = -W_cb L_cb^T over all pivot columns, and the children's entries of the CB columns are
//     added to it (the contribution blocks need no zero fill: every entry of their lower
//     triangles is stored before it is read, and nothing reads above the diagonal).
// Inertia: signs of D (2x2 blocks by their eigenvalues), integer counters. Null pivots: a pivot
// block with an eigenvalue of magnitude <= null_tol * max|K_ij| is flagged (the reference's
// CNTL(3) rule, include/hsolve/factor.hpp:79-115), counted as zero and removed from the
// elimination (like MUMPS' treatment of null pivots).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdlib>
#include <list>
#include <mutex>
#include <stdexcept>
#include <unordered_map>
#include <string>
#include <type_traits>
#include <vector>

#include "engine.hpp"
#include "front_common.hpp"
#include "solve.hpp"

namespace hsg {

constexpr int kThreads = 256;

// ------------------------------------------------------------------------------ small helpers

// Zeroes n elements of p (a device allocation, so 16-byte aligned): 16-byte streaming stores, the
// remainder element-wise.
typedef unsigned int hsg_u32x4 __attribute__((ext_vector_type(4)));
template <class T>
__global__ void k_zero(T* __restrict__ p, int64_t n) {
  constexpr int kV = 16 / sizeof(T);
  const int64_t nv = n / kV;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t first = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  hsg_u32x4* q = reinterpret_cast<hsg_u32x4*>(p);
  const hsg_u32x4 z = {0u, 0u, 0u, 0u};
  for (int64_t i = first; i < nv; i += stride) __builtin_nontemporal_store(z, q + i);
  for (int64_t i = nv * kV + first; i < n; i += stride) p[i] = (T)0;
}

__global__ void k_fill_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Original entries of K = alpha A + beta B into the panels; the largest |K_ij| is reduced into
// the counter slot kCntKnorm (bit pattern of a non-negative double: integer max is value max).
template <class T>
__global__ void k_assemble_original(int64_t nnz, const int64_t* __restrict__ dst, const int64_t* __restrict__ ia,
                                    const int64_t* __restrict__ ib, BatchValues bvals, T* __restrict__ L,
                                    int64_t zL, unsigned long long* cnt) {
  // factorization z = blockIdx.y of a batch: its values, coefficients, panels and counters
  const int z = blockIdx.y;
  const double* __restrict__ av = bvals.a[z];
  const double* __restrict__ bv = bvals.b[z];
  const double alpha = bvals.alpha[z], beta = bvals.beta[z];
  L += z * zL;
  cnt += z * kCntSlots;
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
  // wave reduction, then the workgroup's waves through LDS, then one atomic per workgroup (atomics
  // on the one counter serialize: one per wave of an entry-per-thread grid cost ~0.25 ms)
  __shared__ double wmax[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) local_max = fmax(local_max, __shfl_xor(local_max, off));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = local_max;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mx = wmax[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) mx = fmax(mx, wmax[w]);
    atomicMax(&cnt[kCntKnorm], (unsigned long long)__double_as_longlong(mx));
  }
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
// Rows of the blocked path below the LDS block: column j (original local index) of the block
// continues at base[j * ld + row] for rows [row0, row1) in global memory.
template <class T>
struct BelowRows {
  const T* base = nullptr;
  int64_t ld = 0;
  int row0 = 0, row1 = 0;
};

template <class T>
__device__ double below_max(const BelowRows<T>& b, int j, int lane) {
  double mx = 0.0;
  if (b.base)
    for (int r = b.row0 + lane; r < b.row1; r += 64) mx = fmax(mx, fabs((double)b.base[(int64_t)j * b.ld + r]));
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  return mx;
}

template <class T>
__device__ void bk_factor_lds(T* F, int ld, int ncand, int mrows, int* perm, T* dv, T* dof, int8_t* pt, T* t1,
                              T* t2, double tau, bool detect, PivotCounts& pc, BelowRows<T> below) {
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
      // the rows below an LDS block matter only for a pivot that is itself tiny: scan them lazily
      if (detect && below.base) {
        const double d0 = fabs((double)F[k + k * ld]);
        bool tiny = kind == 1 ? (d0 <= tau && m0 <= tau) : (fmax(m0, m1) <= tau);
        if (tiny) {
          m0 = fmax(m0, below_max(below, perm[k], tid));
          if (kind == 2) m1 = fmax(m1, below_max(below, perm[k + 1], tid));
        }
      }
      if (tid == 0) {
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
  N = N.at(blockIdx.z);
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
  // the front's panel, then each child's update matrix added in: kBatch global loads per thread
  // are issued before the first is used (the loops had waited on every load in turn)
  constexpr int kBatch = 8;
  for (int base = 0; base < m * m; base += kThreads * kBatch) {
    T v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int idx = base + tid + kThreads * u;
      v[u] = (idx < m * m && idx / m < ncol) ? N.L[poff + idx] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int idx = base + tid + kThreads * u;
      if (idx < m * m) F[idx] = v[u];
    }
  }
  __syncthreads();
  for (int q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q) {
    const int ch = S.child_list[q];
    const int cbc = S.sn_m[ch] - S.sn_ncol[ch];
    const int32_t* rel = S.relmap + S.sn_rowoff[ch] + S.sn_ncol[ch];
    const T* cbp = N.CB + S.cb_off[ch];
    for (int base = 0; base < cbc * cbc; base += kThreads * kBatch) {
      T v[kBatch];
      int ri[kBatch], rj[kBatch];
      bool ok[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int idx = base + tid + kThreads * u;
        const int i = idx % cbc, j = idx / cbc;
        ok[u] = idx < cbc * cbc && i >= j;
        v[u] = ok[u] ? cbp[idx] : (T)0;
        ri[u] = ok[u] ? rel[i] : 0;
        rj[u] = ok[u] ? rel[j] : 0;
      }
      // one child's entries land on distinct entries of the front
#pragma unroll
      for (int u = 0; u < kBatch; ++u)
        if (ok[u]) F[ri[u] + rj[u] * m] += v[u];
    }
    __syncthreads();
  }
  const double tau = null_tol * as_double_bits(N.cnt[kCntKnorm]);
  PivotCounts pc{0, 0, 0, 0, 0};
  bk_factor_lds<T>(F, m, ncol, m, perm, dv, dof, pt, t1, t2, tau, detect != 0, pc, BelowRows<T>{});
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

constexpr int kEaCols = 8;   // parent columns per extend-add workgroup
constexpr int kEaBatch = 16; // child entries per thread whose loads are in flight together

// First index i in [0, n) with rel[i] >= key (n if none), rel strictly increasing; called by a
// whole wave, returned to every lane. Each round samples 64 evenly spaced entries and keeps the
// bracket between the last sample below the key and the first one at or above it, so a search
// over n entries takes ceil(log64 n) dependent loads instead of log2 n.
__device__ __forceinline__ int wave_lower_bound(const int32_t* __restrict__ rel, int n, int key, int lane) {
  int lo = 0, len = n;  // the answer lies in [lo, lo + len]
  while (len > 64) {
    const int step = (len + 63) / 64;
    const int off = lane * step;
    const bool less = off < len && rel[lo + off] < key;
    const int c = __popcll(__ballot(less));
    if (c == 0) return lo;
    const int nlo = lo + (c - 1) * step + 1;
    const int nhi = min(lo + c * step, lo + len);
    lo = nlo;
    len = nhi - nlo;
  }
  const bool less = lane < len && rel[lo + lane] < key;
  return lo + __popcll(__ballot(less));
}

// Extend-add into the large fronts of one level: workgroup (x, y) adds, child by child, the
// entries of the children's update matrices that land in kEaCols (8) parent columns of front
// list[y]: part 0 covers the panel columns [kEaCols x, kEaCols x + kEaCols) (before the front is
// factorized), part 1 the columns ncol + [kEaCols x, kEaCols x + kEaCols) of the front's own
// update matrix, after the front's update
// product has been stored there. The child columns that map there are a contiguous range
// (relmap is increasing); the four waves find the ranges of two children at a time, one end
// each. A child's (column, row) pairs of the lower triangle are then spread over the workgroup,
// each thread issuing the loads of kEaBatch entries (row map, child value, target) before using
// any. The entries of one child map to distinct parent entries, and a barrier separates the
// children, so the sums are deterministic and need no atomics.
template <class T>
__global__ void __launch_bounds__(kThreads) k_extend_add_large(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                              int part) {
  N = N.at(blockIdx.z);
  __shared__ int jr[2][2];
  const int s = list[blockIdx.y];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s];
  const int cb = m - ncol;
  const int lim = part == 0 ? ncol : m;
  const int a = (part == 0 ? 0 : ncol) + blockIdx.x * kEaCols;
  if (a >= lim) return;
  const int b = min(lim, a + kEaCols);
  T* P = N.L + S.panel_off[s];
  T* C = N.CB + S.cb_off[s];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  static_assert(kThreads == 256, "extend-add: four waves search two children's ranges at once");
  const int q0 = S.child_ptr[s], q1 = S.child_ptr[s + 1];
  for (int g0 = q0; g0 < q1; g0 += 2) {
    const int ng = min(2, q1 - g0);
    // wave w finds, for child g0 + w / 2, the first column mapping at or beyond a (w even) or b
    if ((wv >> 1) < ng) {
      const int ch = S.child_list[g0 + (wv >> 1)];
      const int32_t* rel = S.relmap + S.sn_rowoff[ch] + S.sn_ncol[ch];
      const int r = wave_lower_bound(rel, S.sn_m[ch] - S.sn_ncol[ch], (wv & 1) ? b : a, lane);
      if (lane == 0) jr[wv >> 1][wv & 1] = r;
    }
    __syncthreads();
    for (int c = 0; c < ng; ++c) {
      const int ch = S.child_list[g0 + c];
      const int cbc = S.sn_m[ch] - S.sn_ncol[ch];
      const int32_t* rel = S.relmap + S.sn_rowoff[ch] + S.sn_ncol[ch];
      const T* cbp = N.CB + S.cb_off[ch];
      const int j0 = jr[c][0], j1 = jr[c][1];
      if (j0 == j1) continue;
      const int R = cbc - j0, nc = j1 - j0;
      const int tot = R * nc;  // pairs (column j0 + t / R, row j0 + t % R); rows above the diagonal skipped
      for (int t0 = 0; t0 < tot; t0 += kThreads * kEaBatch) {
        int ri[kEaBatch], pc[kEaBatch];
        T cv[kEaBatch];
        bool ok[kEaBatch];
  #pragma unroll
        for (int u = 0; u < kEaBatch; ++u) {
          const int t = t0 + tid + kThreads * u;
          const int jj = t / R;
          const int i = j0 + (t - jj * R), j = j0 + jj;
          ok[u] = t < tot && i >= j;
          const int ic = ok[u] ? i : j0, jc = ok[u] ? j : j0;  // the diagonal entry j0 stands in
          ri[u] = rel[ic];
          pc[u] = rel[jc];
          cv[u] = cbp[(int64_t)jc * cbc + ic];
        }
        T* dst[kEaBatch];
        T dv[kEaBatch];
  #pragma unroll
        for (int u = 0; u < kEaBatch; ++u) {
          dst[u] = pc[u] < ncol ? P + (int64_t)pc[u] * m + ri[u] : C + (int64_t)(pc[u] - ncol) * cb + (ri[u] - ncol);
          dv[u] = *dst[u];
        }
  #pragma unroll
        for (int u = 0; u < kEaBatch; ++u)
          if (ok[u]) *dst[u] = dv[u] + cv[u];
      }
      __syncthreads();  // the next child's entries may land on the same parent entries
    }
    __syncthreads();  // every thread has read jr before the next pair's searches rewrite it
  }
}

// Maximum over the 64 lanes of a wave of a non-negative key (bit pattern of a non-negative
// double: integer order = value order), returned to every lane. DPP reduction: row_shr 1, 2, 4,
// 8 leave each 16-lane row's maximum in its lane 15, row_bcast 15 / 31 carry them to lane 63.
// Lanes a DPP step does not write read 0, the neutral element for non-negative keys.
__device__ __forceinline__ unsigned long long wave_max_key(unsigned long long v) {
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
#define HSG_DPP_MAX_STEP(ctrl, rmask)                                                          \
  {                                                                                           \
    const unsigned olo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)lo, ctrl, rmask, 0xf, true); \
    const unsigned ohi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)hi, ctrl, rmask, 0xf, true); \
    const unsigned long long o = ((unsigned long long)ohi << 32) | olo;                      \
    const unsigned long long c = ((unsigned long long)hi << 32) | lo;                        \
    if (o > c) {                                                                              \
      lo = olo;                                                                               \
      hi = ohi;                                                                               \
    }                                                                                         \
  }
  HSG_DPP_MAX_STEP(0x111, 0xf)  // row_shr:1
  HSG_DPP_MAX_STEP(0x112, 0xf)  // row_shr:2
  HSG_DPP_MAX_STEP(0x114, 0xf)  // row_shr:4
  HSG_DPP_MAX_STEP(0x118, 0xf)  // row_shr:8
  HSG_DPP_MAX_STEP(0x142, 0xa)  // row_bcast:15 into rows 1, 3
  HSG_DPP_MAX_STEP(0x143, 0xc)  // row_bcast:31 into rows 2, 3
#undef HSG_DPP_MAX_STEP
  lo = (unsigned)__builtin_amdgcn_readlane((int)lo, 63);
  hi = (unsigned)__builtin_amdgcn_readlane((int)hi, 63);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ double wave_max_abs(double v) {
  return __longlong_as_double((long long)wave_max_key((unsigned long long)__double_as_longlong(fabs(v))));
}

// Maximum of a non-negative v over the 64 lanes and the lowest lane holding it (arg), returned to
// every lane, exactly as wave_max_key + ballot would give, in about a third of the time: the DPP
// reduction runs on the high 32 bits only (a 32-bit max per step, folded into the DPP move), the
// lowest lane whose high word is the maximum is read in full, and a ballot checks that no lane
// with the same high word holds a larger value; only then (equal high words, different low
// words) the full 64-bit reduction runs.
__device__ __forceinline__ double wave_argmax_nonneg(double v, int& arg) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const unsigned hi = (unsigned)(bits >> 32);
  unsigned m = hi;
#define HSG_DPP_UMAX_STEP(ctrl, rmask) \
  m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, ctrl, rmask, 0xf, true));
  HSG_DPP_UMAX_STEP(0x111, 0xf)  // row_shr:1
  HSG_DPP_UMAX_STEP(0x112, 0xf)  // row_shr:2
  HSG_DPP_UMAX_STEP(0x114, 0xf)  // row_shr:4
  HSG_DPP_UMAX_STEP(0x118, 0xf)  // row_shr:8
  HSG_DPP_UMAX_STEP(0x142, 0xa)  // row_bcast:15 into rows 1, 3
  HSG_DPP_UMAX_STEP(0x143, 0xc)  // row_bcast:31 into rows 2, 3
#undef HSG_DPP_UMAX_STEP
  const unsigned mh = (unsigned)__builtin_amdgcn_readlane((int)m, 63);
  const unsigned long long top = __ballot(hi == mh);
  int l0 = __ffsll((long long)top) - 1;
  const unsigned lo0 = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)bits, l0);
  unsigned long long best = ((unsigned long long)mh << 32) | lo0;
  if (__ballot(hi == mh && bits > best)) {
    best = wave_max_key(hi == mh ? bits : 0ull);
    l0 = __ffsll((long long)__ballot(bits == best)) - 1;
  }
  arg = l0;
  return __longlong_as_double((long long)best);
}

// Value of v in lane src (src uniform across the wave).
__device__ __forceinline__ double readlane_of(double v, int src) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, src);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float readlane_of(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// Bunch-Kaufman on the nb x nb diagonal block of panel p of each front, eight 64-lane waves per
// front: lane i of every wave stands for row i of the block (LDS, padded leading dimension).
// Every wave reads the pivot column into a register and runs the same pivot search on it
// (wave_argmax_nonneg: maximum and lowest lane holding it), so all reach the same decision
// without a barrier; the column maxima of the null-pivot test are formed only when a pivot is
// small enough for the test to need them. Wave w then updates the columns [8 w, 8 w + 8) of the trailing
// triangle with the step's multipliers formed from the pivot column (broadcast LDS reads), and
// every lane stores all of its columns (the new value where it changes, the old one elsewhere),
// so the update has no per-lane branches; a wave whose columns all lie left of the pivot skips
// it. The pivot columns of F keep their unscaled values; their multipliers go to a separate
// array Lm, written by wave 0 in the same step (nobody reads Lm during a step, and a stored-back
// old value of F is never stale). Wave 0 alone swaps rows and columns (Lm for the columns
// already pivoted, F for the rest); every lane keeps the same tallies, and lane i the pivot data
// of position i. One barrier per step
// (four when rows are swapped). Eight waves measured faster than four (shorter update per wave,
// and waves drop out sooner) and than sixteen (the redundant searches crowd the SIMDs).
constexpr int kBkLd = 65;
constexpr int kBkWaves = 8;
constexpr int kBkCols = 64 / kBkWaves;  // columns of the block per wave

template <class T>
__global__ void __launch_bounds__(64 * kBkWaves) k_bk_diag(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                           int p, int nb, double null_tol, int detect) {
  N = N.at(blockIdx.z);
  int32_t* __restrict__ bperm = N.bperm;
  __shared__ T F[64 * kBkLd];   // the block; pivot columns keep their unscaled values (and D)
  __shared__ T Lm[64 * kBkLd];  // multipliers of the pivot columns (strictly lower L11)
  __shared__ int perm[64];
  const int f = blockIdx.x;
  const int s = list[f];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int c0 = p * nb;
  if (c0 >= ncol) return;
  const int w = min(nb, ncol - c0);
  T* P = N.L + S.panel_off[s];
  const int lane = threadIdx.x % 64;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const bool lead = wv == 0;
  const int jb = wv * kBkCols;  // this wave updates the block columns [jb, jb + kBkCols)
  {
    // branch-free loads (row index clamped into the block) so that they are issued together
    const T* src = P + (int64_t)c0 * m + c0 + min(lane, w - 1);
#pragma unroll 4
    for (int j = wv; j < w; j += kBkWaves) {
      const T v = src[(int64_t)j * m];
      F[lane + j * kBkLd] = (lane >= j && lane < w) ? v : (T)0;
    }
  }
  if (lead) perm[lane] = lane;
  __syncthreads();
  const double tau = null_tol * as_double_bits(N.cnt[kCntKnorm]);
  const bool det = detect != 0;
  int neg = 0, zero = 0, pos = 0, two = 0, exz = 0;  // tallies (the same in every lane)
  unsigned long long pairs = 0;                      // bit j: column j starts a 2x2 step
  int8_t my_pt = kPiv1;
  T my_dv = (T)0, my_dof = (T)0;
  int k = 0;
  while (k < w) {
    // ---- pivot search (column k, candidates k..w-1), the same in every wave
    const T fk = F[lane + k * kBkLd];
    const double absakk = fabs((double)readlane_of(fk, k));
    const double v = (lane > k && lane < w) ? fabs((double)fk) : 0.0;
    int imax;
    const double colmax = wave_argmax_nonneg(v, imax);
    if (!(colmax > 0.0)) imax = -1;
    int kind = 1, kp = k;
    // the row search reads entries of row imax that this step's update changes: when it runs and no
    // interchange follows, a barrier keeps every wave's reads ahead of the first update store (the
    // condition depends on column k only, which no wave changes in the step, so it is uniform)
    const bool row_search = !(absakk >= kBkAlpha * colmax) && imax > k;
    if (row_search) {
      const T ra = F[imax + lane * kBkLd], rb = F[lane + imax * kBkLd];
      const double r = (lane >= k && lane < w && lane != imax) ? fabs((double)(lane < imax ? ra : rb)) : 0.0;
      int rarg;
      const double rowmax = wave_argmax_nonneg(r, rarg);
      if (absakk * rowmax >= kBkAlpha * colmax * colmax) {
        kind = 1;
        kp = k;
      } else if (fabs((double)F[imax + imax * kBkLd]) >= kBkAlpha * rowmax) {
        kind = 1;
        kp = imax;
      } else {
        kind = 2;
        kp = imax;
      }
    }
    const int kk = kind == 1 ? k : k + 1;
    const bool swapped = kp != kk;
    if (row_search && !swapped) __syncthreads();
    if (swapped) {
      const int pp = kk, q = kp;
      __syncthreads();  // every wave has finished reading the block for its search
      if (lead) {
        if (lane < pp) {  // rows p, q of the columns before p: multipliers, and column k of F when p = k + 1
          T* X = lane < k ? Lm : F;
          const T a = X[pp + lane * kBkLd];
          X[pp + lane * kBkLd] = X[q + lane * kBkLd];
          X[q + lane * kBkLd] = a;
        } else if (lane > pp && lane < q) {  // A(j, p) <-> A(q, j)
          const T a = F[lane + pp * kBkLd];
          F[lane + pp * kBkLd] = F[q + lane * kBkLd];
          F[q + lane * kBkLd] = a;
        } else if (lane > q && lane < w) {  // A(i, p) <-> A(i, q)
          const T a = F[lane + pp * kBkLd];
          F[lane + pp * kBkLd] = F[lane + q * kBkLd];
          F[lane + q * kBkLd] = a;
        }
      }
      __syncthreads();
      if (lead && lane == 0) {
        const T a = F[pp + pp * kBkLd];
        F[pp + pp * kBkLd] = F[q + q * kBkLd];
        F[q + q * kBkLd] = a;
        const int t = perm[pp];
        perm[pp] = perm[q];
        perm[q] = t;
      }
      __syncthreads();
    }
    // pivot column(s) after the interchange, in registers
    const T c1 = swapped ? F[lane + k * kBkLd] : fk;
    const T c2 = kind == 2 ? F[lane + (k + 1) * kBkLd] : (T)0;
    // ---- remaining column maxima for the null-pivot test (rows of the block below the pivot,
    //      then the panel rows below the block), formed only when a pivot is small enough for the
    //      test to need them (the conditions are uniform: every lane reaches the reductions)
    const int s0 = kind == 1 ? k + 1 : k + 2;
    auto max_below = [&]() -> double {
      if (kind == 1 && !swapped) return colmax;  // the same column below the pivot
      int unused;
      double mv = 0.0;
      if (lane >= s0 && lane < w) {
        mv = fabs((double)c1);
        if (kind == 2) mv = fmax(mv, fabs((double)c2));
      }
      return wave_argmax_nonneg(mv, unused);
    };
    const bool own = lead;  // wave 0 writes the multipliers of this step
    T* row = F + lane;
    if (kind == 1) {
      const double d = (double)readlane_of(c1, k);
      bool null_piv = d == 0.0;
      if (det && fabs(d) <= tau && max_below() <= tau) {
        double mb = 0.0;
        const int jc = perm[k];
        for (int r2 = c0 + w + lane; r2 < m; r2 += 64) mb = fmax(mb, fabs((double)P[(int64_t)(c0 + jc) * m + r2]));
        mb = wave_max_abs(mb);
        if (mb <= tau) null_piv = true;
      }
      const T inv = null_piv ? (T)0 : (T)(1.0 / d);
      // tallies and pivot data without lane branches (the tallies are the same in every lane)
      exz += d == 0.0;
      zero += null_piv;
      neg += !null_piv && d < 0;
      pos += !null_piv && !(d < 0);
      my_pt = lane == k ? (null_piv ? kPivNull : kPiv1) : my_pt;
      my_dv = lane == k ? (T)d : my_dv;
      my_dof = lane == k ? (T)0 : my_dof;
      const bool below = lane > k && lane < w;
      const T wi = below ? c1 : (T)0;
      const T li = wi * inv;  // multiplier of row `lane`, zero outside the rows below the pivot
      // rank-1 update of this wave's columns jb + q, skipped once they all lie left of the pivot
      if (jb + kBkCols - 1 > k) {
        // row `lane`: F(lane, j) -= l_lane F(j, k), with F(j, k) read from lane j of the pivot
        // column in registers (a uniform lane index). wi vanishes for j <= k and li outside the
        // rows below the pivot, so the other entries keep their values; where j > lane the entry
        // is in the upper triangle, which nothing reads, so no per-lane mask is needed
        T old[kBkCols];
#pragma unroll
        for (int q = 0; q < kBkCols; ++q) old[q] = row[(jb + q) * kBkLd];
#pragma unroll
        for (int q = 0; q < kBkCols; ++q) row[(jb + q) * kBkLd] = old[q] - li * readlane_of(wi, jb + q);
      }
      if (own) Lm[lane + k * kBkLd] = li;  // zero above the rows below the pivot: never read
      __syncthreads();
      k += 1;
    } else {
      const double a = (double)readlane_of(c1, k), b = (double)readlane_of(c1, k + 1),
                   c = (double)readlane_of(c2, k + 1);
      const double dt = a * c - b * b;
      const double mean = 0.5 * (a + c), rad = sqrt(0.25 * (a - c) * (a - c) + b * b);
      const double e1 = mean - rad, e2 = mean + rad;
      // coupled: the pair's columns have entries above tau below the pair (then a small
      // eigenvalue of the pair is not a null pivot)
      const bool maybe_null = det && (fabs(e1) <= tau || fabs(e2) <= tau);
      const bool coupled = maybe_null && max_below() > tau;
      int nz = 0, nn = 0, np = 0;
      for (double e : {e1, e2}) {
        const bool z = (det && fabs(e) <= tau && !coupled) || e == 0.0;
        if (z) ++nz;
        else if (e < 0) ++nn;
        else ++np;
      }
      const bool dead = nz > 0 || dt == 0.0;
      const T i11 = dead ? (T)0 : (T)(c / dt), i12 = dead ? (T)0 : (T)(-b / dt), i22 = dead ? (T)0 : (T)(a / dt);
      exz += dt == 0.0;
      zero += nz;
      neg += nn;
      pos += np;
      ++two;
      pairs |= 1ull << k;
      my_pt = lane == k ? (dead ? kPivNull : kPiv2First) : (lane == k + 1 ? (dead ? kPivNull : kPiv2Second) : my_pt);
      my_dv = lane == k ? (T)a : (lane == k + 1 ? (T)c : my_dv);
      my_dof = lane == k ? (T)b : (lane == k + 1 ? (T)0 : my_dof);
      const bool below = lane > k + 1 && lane < w;
      const T w1 = below ? c1 : (T)0;
      const T w2 = below ? c2 : (T)0;
      // rank-2 update of this wave's columns, unmasked as in the 1x1 step (l1v, l2v vanish for
      // j <= k + 1 and when dead), skipped once they all lie left of the pivot pair
      if (jb + kBkCols - 1 > k + 1) {
        T l1v[kBkCols], l2v[kBkCols];
#pragma unroll
        for (int q = 0; q < kBkCols; ++q) {
          const int j = jb + q;
          const bool in = j > k + 1 && j < w;
          const T f1 = F[j + k * kBkLd], f2 = F[j + (k + 1) * kBkLd];
          const T a1 = in ? f1 : (T)0, a2 = in ? f2 : (T)0;
          l1v[q] = a1 * i11 + a2 * i12;
          l2v[q] = a1 * i12 + a2 * i22;
        }
        T old[kBkCols];
#pragma unroll
        for (int q = 0; q < kBkCols; ++q) old[q] = row[(jb + q) * kBkLd];
#pragma unroll
        for (int q = 0; q < kBkCols; ++q) row[(jb + q) * kBkLd] = old[q] - (w1 * l1v[q] + w2 * l2v[q]);
      }
      if (own) {  // zero above the rows below the pair: never read
        Lm[lane + k * kBkLd] = w1 * i11 + w2 * i12;
        Lm[lane + (k + 1) * kBkLd] = w1 * i12 + w2 * i22;
      }
      __syncthreads();
      k += 2;
    }
  }
  __syncthreads();
  // ---- write back L11 and D, pivot data
  for (int j = wv; j < w; j += kBkWaves)
    if (lane >= j && lane < w)
      P[(int64_t)(c0 + j) * m + c0 + lane] =
          (lane == j || (lane == j + 1 && ((pairs >> j) & 1))) ? F[lane + j * kBkLd] : Lm[lane + j * kBkLd];
  if (lead && lane < w) {
    N.dval[first + c0 + lane] = my_dv;
    N.doff[first + c0 + lane] = my_dof;
    N.ptype[first + c0 + lane] = my_pt;
    N.piv[first + c0 + lane] = c0 + perm[lane];
    N.ipiv[first + c0 + perm[lane]] = c0 + lane;
    bperm[(int64_t)f * nb + lane] = perm[lane];
  }
  if (lead && lane == 0) {
    unsigned long long* cnt = N.cnt;
    if (neg) atomicAdd(&cnt[kCntNeg], (unsigned long long)neg);
    if (zero) atomicAdd(&cnt[kCntZero], (unsigned long long)zero);
    if (pos) atomicAdd(&cnt[kCntPos], (unsigned long long)pos);
    if (two) atomicAdd(&cnt[kCnt2x2], (unsigned long long)two);
    if (exz) atomicAdd(&cnt[kCntExactZero], (unsigned long long)exz);
  }
}

// Bunch-Kaufman on the nb x nb diagonal block of panel p of each front with one 64-lane wave per
// front and no barrier between pivot steps (a workgroup of one wave: the LDS fences are waits on
// its own LDS operations). Lane i keeps row i of the symmetric block in registers, both triangles,
// and the two copies of every entry stay bitwise equal: each step subtracts from entry (i, j) the
// product of two per-lane factors u_i u_j times a sign, rounded once by one FMA, which is the same
// in row i and row j. For a 1x1 pivot d, u_i = c_i / sqrt|d| (c: the pivot column); for a 2x2 pivot
// D = J diag(l1, l2) J^T (one Jacobi rotation, Golub and Van Loan's sym.schur2) and
// [u1_i, u2_i] = diag(1 / sqrt|l|) J^T [c1_i, c2_i], two FMAs. Lane i = 16 r + t holds column j of
// its row in register (j - 16 r) mod 64, so that the update needs no LDS: the factors are rotated by
// 16, 32 and 48 lanes (ds_bpermute) and each of the 64 FMAs takes its factor u_j from lane n of its
// own 16-lane row of one rotation (64-bit DPP row_newbcast). Interchanges only exchange the logical
// positions of two lanes, so no data moves; the pivot column is read out of every lane's row at the
// uniform offset j mod 16 of its four quarters and a per-lane choice of the rotation.
// The diagonal is kept in a separate register, updated with the same FMA as its row copy. The
// pivot search reduces over the lanes below the pivot position (ties to the lowest position, as
// with physical interchanges). The multipliers go to LDS by physical row and logical column and
// are written to the panel in logical order at the end, with D, the pivot types and the block
// permutation, in the same layout as k_bk_diag. The pivot rules (alpha, the row search, the
// null-pivot tests and their tallies) are those of k_bk_diag.
template <class T>
__device__ __forceinline__ double wave_max_abs_masked(T v, bool on) {
  int unused;
  return wave_argmax_nonneg(on ? fabs((double)v) : 0.0, unused);
}

// 1 / sqrt(x) for x > 0 to full double precision: the hardware estimate and two Newton steps
// y <- y + y (1 - x y^2) / 2.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double e = fma(-(x * y), y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}
// 1 / x to full double precision: the hardware estimate and two Newton steps y <- y + y (1 - x y).
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int i = 0; i < 2; ++i) y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

// acc[n] += b * (src of lane n of this lane's 16-lane row), n = 0..15: v_fmac_f64 with the
// 64-bit DPP row broadcast. The s_nop covers the wait states between a VALU write of src and its
// DPP read.
__device__ __forceinline__ void fmac_row_bcast16(double (&acc)[16], double src, double b) {
  __asm__("s_nop 1\n"
          "v_fmac_f64_dpp %0, %16, %17 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %1, %16, %17 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %2, %16, %17 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %3, %16, %17 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %4, %16, %17 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %5, %16, %17 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %6, %16, %17 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %7, %16, %17 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %8, %16, %17 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %9, %16, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %10, %16, %17 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %11, %16, %17 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %12, %16, %17 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %13, %16, %17 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %14, %16, %17 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
          "v_fmac_f64_dpp %15, %16, %17 row_newbcast:15 row_mask:0xf bank_mask:0xf"
          : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]),
            "+v"(acc[7]), "+v"(acc[8]), "+v"(acc[9]), "+v"(acc[10]), "+v"(acc[11]), "+v"(acc[12]), "+v"(acc[13]),
            "+v"(acc[14]), "+v"(acc[15])
          : "v"(src), "v"(b));
}

// An upper bound of the maximum of a non-negative v over the 64 lanes, returned to every lane: the
// maximum of the high words by DPP with all low-word bits set (at least every lane's value; a NaN,
// which fails every comparison, when the maximum is infinite).
__device__ __forceinline__ double wave_max_nonneg_bound(double v) {
  unsigned m = (unsigned)((unsigned long long)__double_as_longlong(v) >> 32);
#define HSG_DPP_UMAX_STEP(ctrl, rmask) \
  m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, ctrl, rmask, 0xf, true));
  HSG_DPP_UMAX_STEP(0x111, 0xf)  // row_shr:1
  HSG_DPP_UMAX_STEP(0x112, 0xf)  // row_shr:2
  HSG_DPP_UMAX_STEP(0x114, 0xf)  // row_shr:4
  HSG_DPP_UMAX_STEP(0x118, 0xf)  // row_shr:8
  HSG_DPP_UMAX_STEP(0x142, 0xa)  // row_bcast:15 into rows 1, 3
  HSG_DPP_UMAX_STEP(0x143, 0xc)  // row_bcast:31 into rows 2, 3
#undef HSG_DPP_UMAX_STEP
  const unsigned mh = (unsigned)__builtin_amdgcn_readlane((int)m, 63);
  return __longlong_as_double((long long)(((unsigned long long)mh << 32) | 0xffffffffull));
}

// v of lane (lane + 16 s) mod 64
__device__ __forceinline__ double rotate_lanes(double v, int lane, int s) {
  const int addr = ((lane + 16 * s) & 63) * 4;
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)bits);
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(bits >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Maximum of a non-negative v over the 64 lanes, returned to every lane: the maximum of the high
// words by DPP (v_max_u32 with row_shr 1, 2, 4, 8 and row_bcast 15 / 31), the low word of the
// lowest lane holding it, and a ballot that no lane with the same high word holds a larger value
// (else the full 64-bit reduction).
__device__ __forceinline__ double wave_max_nonneg(double v) {
  int unused;
  return wave_argmax_nonneg(v, unused);
}

constexpr int kBkWaveLoaders = 8;  // waves of k_bk_wave that load and store the block

__global__ void __launch_bounds__(64 * kBkWaveLoaders) k_bk_wave(DevSym S, DevNum<double> N,
                                                                 const int32_t* __restrict__ list, int p, int nb,
                                                                 double null_tol, int detect) {
  typedef double T;
  N = N.at(blockIdx.z);
  __shared__ T F[64 * kBkLd];  // the block's lower triangle, then the multipliers [physical row][logical column]
  __shared__ T sdof[64];
  const int f = blockIdx.x;
  const int s = list[f];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int c0 = p * nb;
  if (c0 >= ncol) return;
  const int w = min(nb, ncol - c0);
  T* P = N.L + S.panel_off[s];
  const int lane = threadIdx.x % 64;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  {
    // the lower triangle by columns, kBkWaveLoaders columns at a time (loads issued before stores)
    const T* src = P + (int64_t)c0 * m + c0 + min(lane, w - 1);
    T v[64 / kBkWaveLoaders];
#pragma unroll
    for (int t = 0; t < 64 / kBkWaveLoaders; ++t) v[t] = src[(int64_t)min(wv + kBkWaveLoaders * t, w - 1) * m];
#pragma unroll
    for (int t = 0; t < 64 / kBkWaveLoaders; ++t) {
      const int j = wv + kBkWaveLoaders * t;
      F[lane + j * kBkLd] = (lane >= j && lane < w && j < w) ? v[t] : (T)0;
    }
  }
  __syncthreads();
  // the other waves end here and give their registers back to the CU (a workgroup of eight waves
  // at this register count would fill it); the pivot steps and the write-back are wave 0's
  if (wv != 0) return;
  const int rr = lane >> 4;  // this lane's 16-lane row: register R holds column (R + 16 rr) mod 64
  typedef T Quarter __attribute__((ext_vector_type(16)));
  Quarter q0, q1, q2, q3;  // registers 16 s .. 16 s + 15 of the row in qs
  auto load_quarter = [&](Quarter& q, int sq) {
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const int j = (16 * (sq + rr) + n) & 63;
      q[n] = lane >= j ? F[lane + j * kBkLd] : F[j + lane * kBkLd];
    }
  };
  load_quarter(q0, 0);
  load_quarter(q1, 1);
  load_quarter(q2, 2);
  load_quarter(q3, 3);
  T col = F[lane];  // column 0 (lane 0's row): the first pivot column
  // F holds the multipliers from here on (this wave's LDS operations complete in order)
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  // lanes of 16-lane row r (a uniform 64-bit mask, used as the lane condition of a select)
  auto row_lanes = [](int r) -> unsigned long long { return 0xffffull << (16 * (r & 3)); };
  // column j of the row (j uniform): register (j - 16 rr) mod 64, read with a uniform branch on
  // j mod 16 (the quarters live in separately allocated registers, which a register-indexed read
  // cannot address); the lanes of row r take it from quarter (j / 16 - r) mod 4
  auto pick = [&](int j) -> T {
    j = __builtin_amdgcn_readfirstlane(j);
    T a, b, c, d;
    switch (j & 15) {
#define HSG_BK_PICK_CASE(n) \
  case n:                   \
    a = q0[n];              \
    b = q1[n];              \
    c = q2[n];              \
    d = q3[n];              \
    break;
      HSG_BK_PICK_CASE(0) HSG_BK_PICK_CASE(1) HSG_BK_PICK_CASE(2) HSG_BK_PICK_CASE(3)
      HSG_BK_PICK_CASE(4) HSG_BK_PICK_CASE(5) HSG_BK_PICK_CASE(6) HSG_BK_PICK_CASE(7)
      HSG_BK_PICK_CASE(8) HSG_BK_PICK_CASE(9) HSG_BK_PICK_CASE(10) HSG_BK_PICK_CASE(11)
      HSG_BK_PICK_CASE(12) HSG_BK_PICK_CASE(13) HSG_BK_PICK_CASE(14) default: HSG_BK_PICK_CASE(15)
#undef HSG_BK_PICK_CASE
    }
    const int hi = j >> 4;
    T v = __builtin_amdgcn_inverse_ballot_w64(row_lanes(hi + 2)) ? c : d;
    v = __builtin_amdgcn_inverse_ballot_w64(row_lanes(hi + 3)) ? b : v;
    return __builtin_amdgcn_inverse_ballot_w64(row_lanes(hi)) ? a : v;
  };
  // quarter q += b * u_j over its columns j (src: u rotated so that lane n of each 16-lane row
  // holds the factor of the column of register n)
  auto apply_quarter = [&](Quarter& q, T src, T b) {
    T acc[16];
#pragma unroll
    for (int n = 0; n < 16; ++n) acc[n] = q[n];
    fmac_row_bcast16(acc, src, b);
#pragma unroll
    for (int n = 0; n < 16; ++n) q[n] = acc[n];
  };
  const double tau = null_tol * as_double_bits(N.cnt[kCntKnorm]);
  const bool det = detect != 0;
  const unsigned long long valid = w == 64 ? ~0ull : ((1ull << w) - 1);  // lanes of the block's rows
  unsigned long long done = 0;   // lanes of the positions before k
  unsigned long long pairs = 0;  // bit j: position j starts a 2x2 step
  int at = lane;                 // this lane's logical position
  int lop = lane;                // the lane at logical position `lane`
  // pivot data of position `lane`: D entries, pivot type, and the step's inertia counts (neg,
  // zero, pos: 2 bits each; bit 6: an exactly zero pivot)
  T my_dv = 0, my_dof = 0;
  int my_pt = kPiv1, my_code = 0;
  // the previous step's update, applied to the row during the next step (quarter 0 first, the
  // others after the column maximum): one term (u1, b1) after a 1x1 pivot, two (u2, b2
  // first) after a 2x2 pivot; `col` already includes it. x, y, z: u rotated by 16, 32, 48 lanes.
  int npend = 0;
  T u1 = 0, b1 = 0, u2 = 0, b2 = 0;
  T x1 = 0, y1 = 0, z1 = 0, x2 = 0, y2 = 0, z2 = 0;
  auto apply_rest = [&]() {
    if (npend == 2) {
      apply_quarter(q1, x2, b2);
      apply_quarter(q2, y2, b2);
      apply_quarter(q3, z2, b2);
    }
    if (npend >= 1) {
      apply_quarter(q1, x1, b1);
      apply_quarter(q2, y1, b1);
      apply_quarter(q3, z1, b1);
    }
  };
  int k = 0, pk = 0;  // pk: the lane at position k
  while (k < w) {
    if (npend == 2) apply_quarter(q0, u2, b2);
    if (npend >= 1) apply_quarter(q0, u1, b1);
    // ---- column k (exact: the pending update is included), its diagonal and its maximum below
    const unsigned long long bmask = valid & ~done & ~(1ull << pk);  // lanes of the rows below position k
    const bool below = __builtin_amdgcn_inverse_ballot_w64(bmask);
    const T fk = col;
    const double dk = readlane_of(fk, pk);
    const double absakk = fabs(dk);
    const double cbound = wave_max_nonneg_bound(below ? fabs(fk) : 0.0);  // >= the column maximum
    const bool mine = __builtin_amdgcn_inverse_ballot_w64(1ull << k);  // lane k keeps position k's data
    apply_rest();  // (after the reduction above, which gives the rotated factors time to arrive)
    int kn;
    T nu1, nb1, nu2 = 0, nb2 = 0;  // this step's update
    int nterms;
    if (__builtin_expect(absakk >= kBkAlpha * cbound && dk != 0.0 && !(det && absakk <= tau), 1)) {
      // ---- the common step: a 1x1 pivot in place (it passes the test against the exact column
      //      maximum too), not small enough for the null-pivot test
      const double y = rsqrt_nr(absakk);  // 1 / sqrt|d|
      const double inv = dk < 0 ? -(y * y) : y * y;
      my_dv = mine ? dk : my_dv;
      my_dof = mine ? 0.0 : my_dof;
      my_pt = mine ? (int)kPiv1 : my_pt;
      my_code = mine ? (dk < 0 ? 1 : 16) : my_code;
      F[lane + k * kBkLd] = below ? fk * inv : 0.0;  // multiplier of this row (zero outside the rows below)
      nu1 = below ? fk * y : 0.0;
      nb1 = dk < 0 ? nu1 : -nu1;
      nterms = 1;
      done |= 1ull << pk;
      kn = k + 1;
    } else {
      // ---- the general step: the row search, interchanges, 2x2 pivots, null pivots
      const double colmax = wave_max_nonneg(below ? fabs(fk) : 0.0);
      // the lane of the maximum below the pivot; ties: the lowest position among the lanes holding it
      const unsigned long long hold = __ballot(below && fabs(fk) == colmax);
      int imax = __ffsll((long long)hold) - 1;
      if (colmax > 0.0 && __popcll(hold) > 1) {
        int bp = 1 << 30;
        for (unsigned long long h = hold; h; h &= h - 1) {
          const int l = __ffsll((long long)h) - 1;
          const int pl = __builtin_amdgcn_readlane(at, l);
          if (pl < bp) {
            bp = pl;
            imax = l;
          }
        }
      }
      if (!(colmax > 0.0)) imax = -1;
      int kind = 1;
      bool swapped = false;
      T ri = 0;  // column imax, when the row search reads it
      if (!(absakk >= kBkAlpha * colmax) && imax >= 0) {
        ri = pick(imax);
        const bool rs = __builtin_amdgcn_inverse_ballot_w64(valid & ~done & ~(1ull << imax));
        const double rowmax = wave_max_nonneg(rs ? fabs(ri) : 0.0);
        if (absakk * rowmax >= kBkAlpha * colmax * colmax) {
          kind = 1;
        } else if (fabs(readlane_of(ri, imax)) >= kBkAlpha * rowmax) {
          kind = 1;
          swapped = true;
        } else {
          kind = 2;
          swapped = __builtin_amdgcn_readlane(at, imax) != k + 1;
        }
      }
      if (swapped) {
        // positions kk and q = at(imax) exchange their lanes lkk and imax
        const int kk = kind == 1 ? k : k + 1;
        const int q = __builtin_amdgcn_readlane(at, imax);
        const int lkk = __builtin_amdgcn_readlane(lop, kk);
        at = lane == lkk ? q : (lane == imax ? kk : at);
        lop = lane == kk ? imax : (lane == q ? lkk : lop);
      }
      kn = k + kind;
      if (kind == 1) {
        const int piv = swapped ? imax : pk;
        const T c = swapped ? ri : fk;
        const double d = swapped ? readlane_of(ri, imax) : dk;
        done |= 1ull << piv;
        const bool low = __builtin_amdgcn_inverse_ballot_w64(valid & ~done);  // rows below, after the interchange
        bool null_piv = d == 0.0;
        if (det && fabs(d) <= tau) {
          const double mbl = swapped ? wave_max_nonneg(low ? fabs(c) : 0.0) : colmax;
          if (mbl <= tau) {
            double mb = 0.0;
            for (int r2 = c0 + w + lane; r2 < m; r2 += 64) mb = fmax(mb, fabs(P[(int64_t)(c0 + piv) * m + r2]));
            mb = wave_max_abs(mb);
            if (mb <= tau) null_piv = true;
          }
        }
        const double y = null_piv ? 0.0 : rsqrt_nr(fabs(d));  // 1 / sqrt|d|
        const double inv = d < 0 ? -(y * y) : y * y;
        my_dv = mine ? d : my_dv;
        my_dof = mine ? 0.0 : my_dof;
        my_pt = mine ? (null_piv ? (int)kPivNull : (int)kPiv1) : my_pt;
        my_code = mine ? (null_piv ? (4 | (d == 0.0 ? 64 : 0)) : (d < 0 ? 1 : 16)) : my_code;
        const T ci = low ? c : 0.0;
        F[lane + k * kBkLd] = ci * inv;
        nu1 = ci * y;
        nb1 = d < 0 ? nu1 : -nu1;
        nterms = 1;
      } else {
        const int pv2 = imax;  // at position k + 1 now
        const T c1 = fk, c2 = ri;
        const double a = dk, b = readlane_of(c1, pv2), cc = readlane_of(c2, pv2);
        const double dt = a * cc - b * b;
        const double mean = 0.5 * (a + cc), rad = sqrt(0.25 * (a - cc) * (a - cc) + b * b);
        const double e1 = mean - rad, e2 = mean + rad;
        done |= (1ull << pk) | (1ull << pv2);
        const bool low = __builtin_amdgcn_inverse_ballot_w64(valid & ~done);
        const bool maybe_null = det && (fabs(e1) <= tau || fabs(e2) <= tau);
        bool coupled = false;
        if (maybe_null) coupled = wave_max_abs(low ? fmax(fabs(c1), fabs(c2)) : 0.0) > tau;
        int code = 0;
        for (double e : {e1, e2}) {
          const bool z = (det && fabs(e) <= tau && !coupled) || e == 0.0;
          code += z ? 4 : (e < 0 ? 1 : 16);
        }
        const bool dead = (code & 12) != 0 || dt == 0.0;
        const double rdt = dead ? 0.0 : rcp_nr(dt);
        const double i11 = cc * rdt, i12 = -b * rdt, i22 = a * rdt;
        pairs |= 1ull << k;
        const bool next = __builtin_amdgcn_inverse_ballot_w64(1ull << (k + 1));
        my_dv = mine ? a : (next ? cc : my_dv);
        my_dof = mine ? b : (next ? 0.0 : my_dof);
        my_pt = mine ? (dead ? (int)kPivNull : (int)kPiv2First) : (next ? (dead ? (int)kPivNull : (int)kPiv2Second) : my_pt);
        my_code = mine ? (code | (dt == 0.0 ? 64 : 0)) : (next ? 0 : my_code);
        const T w1 = low ? c1 : 0.0, w2 = low ? c2 : 0.0;
        F[lane + k * kBkLd] = w1 * i11 + w2 * i12;
        F[lane + (k + 1) * kBkLd] = w1 * i12 + w2 * i22;
        // D = J diag(l1, l2) J^T, J = [cs sn; -sn cs] (uniform values; b = 0: no rotation, the
        // selects keep the step free of branches)
        const bool rot = b != 0.0;
        const double tq = (cc - a) * (0.5 * rcp_nr(rot ? b : 1.0));
        const double h = fma(tq, tq, 1.0);
        const double t = rot ? (tq >= 0.0 ? 1.0 : -1.0) * rcp_nr(fabs(tq) + h * rsqrt_nr(h)) : 0.0;
        const double cs = rsqrt_nr(fma(t, t, 1.0)), sn = t * cs;
        const double l1 = a - t * b, l2 = cc + t * b;
        const bool upd = !dead && l1 != 0.0 && l2 != 0.0;
        const double s1 = upd ? rsqrt_nr(fabs(l1)) : 0.0, s2 = upd ? rsqrt_nr(fabs(l2)) : 0.0;
        nu1 = (cs * w1 - sn * w2) * s1;
        nu2 = (sn * w1 + cs * w2) * s2;
        nb1 = l1 < 0.0 ? nu1 : -nu1;
        nb2 = l2 < 0.0 ? nu2 : -nu2;
        nterms = 2;
      }
    }
    npend = nterms;
    u1 = nu1;
    b1 = nb1;
    u2 = nu2;
    b2 = nb2;
    if (kn < w) {
      // the next pivot column: the row's column of the lane at position kn, with this step's update
      // formed by the same FMAs that the row gets during the next step
      pk = __builtin_amdgcn_readlane(lop, kn);
      T cn = pick(pk);
      if (npend == 2) cn = fma(readlane_of(u2, pk), b2, cn);
      col = fma(readlane_of(u1, pk), b1, cn);
      x1 = rotate_lanes(u1, lane, 1);
      y1 = rotate_lanes(u1, lane, 2);
      z1 = rotate_lanes(u1, lane, 3);
      if (npend == 2) {
        x2 = rotate_lanes(u2, lane, 1);
        y2 = rotate_lanes(u2, lane, 2);
        z2 = rotate_lanes(u2, lane, 3);
      }
    }
    k = kn;
  }
  if (lane < w) {
    N.dval[first + c0 + lane] = my_dv;
    N.doff[first + c0 + lane] = my_dof;
    N.ptype[first + c0 + lane] = (int8_t)my_pt;
  }
  sdof[lane] = my_dof;
  // tallies: the steps' counts summed over the positions (bit b of each 2-bit count by ballot)
  const int code = lane < w ? my_code : 0;
  auto count2 = [&](int sh) -> int {
    return __popcll(__ballot((code >> sh) & 1)) + 2 * __popcll(__ballot((code >> (sh + 1)) & 1));
  };
  const int neg = count2(0), zero = count2(2), pos = count2(4);
  const int exz = __popcll(__ballot((code >> 6) & 1));
  const int two = __popcll(pairs);
  if (lane == 0) {
    unsigned long long* cnt = N.cnt;
    if (neg) atomicAdd(&cnt[kCntNeg], (unsigned long long)neg);
    if (zero) atomicAdd(&cnt[kCntZero], (unsigned long long)zero);
    if (pos) atomicAdd(&cnt[kCntPos], (unsigned long long)pos);
    if (two) atomicAdd(&cnt[kCnt2x2], (unsigned long long)two);
    if (exz) atomicAdd(&cnt[kCntExactZero], (unsigned long long)exz);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  // ---- write back L11 and D in logical order, the permutation: the multipliers of row `lane`
  //      (columns j < lane), then its diagonal entry and the coupling entry of a 2x2 pivot
  const int ph = lop;  // the physical row at logical position `lane`
  T* prow = P + (int64_t)c0 * m + c0 + lane;
  const T* frow = F + ph;
#pragma unroll 8
  for (int j = 0; j < w; ++j)
    if (lane > j && lane < w) prow[(int64_t)j * m] = frow[j * kBkLd];
  if (lane < w) {
    prow[(int64_t)lane * m] = my_dv;
    if (lane > 0 && ((pairs >> (lane - 1)) & 1)) prow[(int64_t)(lane - 1) * m] = sdof[lane - 1];
    N.piv[first + c0 + lane] = c0 + ph;
    N.ipiv[first + c0 + ph] = c0 + lane;
    N.bperm[(int64_t)f * nb + lane] = ph;
  }
}

constexpr int kPanelRows = 64;  // rows per panel-solve workgroup
constexpr int kLqLd = 18;       // padded stride of the panel solve's L11 layout (conflict-free quads)

// v of lane S of each quad of lanes, in every lane of the quad (DPP quad_perm [S, S, S, S]).
template <int S>
__device__ __forceinline__ int quad_bcast_bits(int v) {
  return __builtin_amdgcn_mov_dpp(v, S * 0x55, 0xf, 0xf, false);
}
template <int S>
__device__ __forceinline__ double quad_bcast_c(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = quad_bcast_bits<S>((int)bits), hi = quad_bcast_bits<S>((int)(bits >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int S>
__device__ __forceinline__ float quad_bcast_c(float v) {
  return __int_as_float(quad_bcast_bits<S>(__float_as_int(v)));
}
// s is a compile-time constant wherever the callers' loops are unrolled
template <class T>
__device__ __forceinline__ T quad_bcast(T v, int s) {
  switch (s) {
    case 0: return quad_bcast_c<0>(v);
    case 1: return quad_bcast_c<1>(v);
    case 2: return quad_bcast_c<2>(v);
    default: return quad_bcast_c<3>(v);
  }
}

// Panel solve below the diagonal block: for rows r in [c0 + w, m) of the panel columns
// [c0, c0 + w): apply the block's column permutation, W = A21 L11^-T by forward substitution with
// the block's L11 (W is stored in the W workspace), L21 = W D^-1 (stored in the panel). Workgroups beyond the row chunks permute rows
// [c0, c0 + w) of the earlier panel columns [0, c0).
template <class T>
__global__ void __launch_bounds__(kThreads) k_panel_apply(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                         int p, int nb) {
  N = N.at(blockIdx.z);
  const int32_t* __restrict__ bperm = N.bperm;
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
  // L11 by column and lane of a quad: Lq[(4 j + g) * kLqLd + u] = L11(g + 4 u, j), zero outside
  // the strictly lower triangle and at the (j + 1, j) slot of a 2x2 or null pivot pair, which
  // holds D's coupling, not L
  T* Lq = reinterpret_cast<T*>(smem_raw);
  T* X = Lq + 64 * 4 * kLqLd;  // 64 x 64, ld kPanelRows; zero outside rn x w
  __shared__ T dinv[3 * 64];
  __shared__ int8_t ptl[64];
  __shared__ int32_t sbp[64];
  if (tid < w) {
    const int j = tid;
    const int8_t t = N.ptype[first + c0 + j];
    ptl[j] = t;
    sbp[j] = bp[j];
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
  __syncthreads();
  // stage L11 and the permuted columns of A21: row li, columns lj + 4 q. All 32 loads are issued
  // before the first is used (addresses clamped into the block, values masked afterwards).
  static_assert(kThreads == 256 && kPanelRows == 64, "panel solve layout: 64 rows, four lanes per row");
  const int li = tid & 63, lj = tid >> 6;
  {
    T lv[16], av[16];
    const int lic = min(li, w - 1), ric = min(li, rn - 1);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int jc = min(lj + 4 * q, w - 1);
      lv[q] = P[(int64_t)(c0 + jc) * m + c0 + lic];
      av[q] = P[(int64_t)(c0 + sbp[jc]) * m + r0 + ric];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int j = lj + 4 * q;
      const bool jin = j < w;
      const bool coupling = jin && li == j + 1 && (ptl[j] == kPiv2First || ptl[j] == kPivNull);
      Lq[(4 * j + (li & 3)) * kLqLd + (li >> 2)] = (jin && li > j && li < w && !coupling) ? lv[q] : (T)0;
      X[li + j * kPanelRows] = (jin && li < rn) ? av[q] : (T)0;
    }
  }
  __syncthreads();
  // W = A21 L11^-T by forward substitution along each row, four lanes per row: lane g of the
  // row's quad holds W(r, g + 4 u) in registers. At step j the quad takes W(r, j) from the lane
  // that owns it (DPP broadcast within the quad) and every lane subtracts L11(i, j) W(r, j) from
  // its entries i; those updates are independent, so the dependent chain is one broadcast and
  // one FMA per column. Entries i <= j meet zeros of Lq and keep their values.
  {
    const int rr = tid >> 2, g = tid & 3;
    T xv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) xv[u] = X[rr + (g + 4 * u) * kPanelRows];
    const T* Lg = Lq + g * kLqLd;
#pragma unroll
    for (int j = 0; j < 64; ++j) {
      if (j < w) {
        const T wj = quad_bcast(xv[j >> 2], j & 3);
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (4 * u + 3 > j) xv[u] -= Lg[4 * j * kLqLd + u] * wj;
      }
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) X[rr + (g + 4 * u) * kPanelRows] = xv[u];
  }
  __syncthreads();
  T* Wp = N.W + S.w_off[s];
  if (li < rn) {
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
      const int j = lj + 4 * q;
      if (j >= w) break;
      const T x = X[li + j * kPanelRows];
      Wp[(int64_t)(c0 + j) * m + r0 + li] = x;
      // L21 = W D^-1
      const int8_t t = ptl[j];
      T v;
      if (t == kPiv1) {
        v = x * dinv[3 * j];
      } else if (t == kPiv2First) {
        v = x * dinv[3 * j] + X[li + (j + 1) * kPanelRows] * dinv[3 * j + 1];
      } else if (t == kPiv2Second) {
        v = X[li + (j - 1) * kPanelRows] * dinv[3 * (j - 1) + 1] + x * dinv[3 * (j - 1) + 2];
      } else {
        v = (T)0;
      }
      P[(int64_t)(c0 + j) * m + r0 + li] = v;
    }
  }
}

// C -= A B^T restricted to the lower triangle, per front:
//   mode 0 (trailing update of panel p): rows [c1, m) x cols [c1, ncol) of the panel,
//          A = W(:, c0:c1), B = L(:, c0:c1);
//   mode 1 (contribution block): rows/cols [ncol, m) into CB, A = W(:, 0:ncol), B = L(:, 0:ncol).
// 64 x 64 output tiles, K staged through LDS in steps of 16. The products run on the matrix cores:
// each of the four waves owns a 32 x 32 quarter of the tile as 2 x 2 blocks of 16 x 16, one
// V_MFMA_*_16X16X4 per block and 4-wide K slice (FP64: v_mfma_f64_16x16x4_f64, FP32:
// v_mfma_f32_16x16x4_f32, accumulation in the factor's precision). The operands are swapped
// (the matrix core's row operand is the L side, its column operand the W side) so that a lane's
// accumulator column indexes a tile row: lane l holds row 16 b + l % 16 of four columns, and the
// 16 lanes of one column store 128 contiguous bytes. Which four: accumulator register v of lane l
// is output row 4 (l / 16) + v of the FP32 instruction but row 4 v + l / 16 of the FP64 one
// (MatrixCore::row; the same register layout composable_kernel's mfma_type tables describe as
// group_size 4 x 1 group for f32_16x16x4 and group_size 1 x 4 groups for f64_16x16x4).
constexpr int kTile = 64;
constexpr int kKStep = 16;

template <class T>
struct MatrixCore;
template <>
struct MatrixCore<double> {
  typedef double acc4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc4 fma(double a, double b, acc4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output row (of the 16 x 16 block) of accumulator register v in lane group g = lane / 16
  static __device__ __forceinline__ int row(int g, int v) { return 4 * v + g; }
};
template <>
struct MatrixCore<float> {
  typedef float acc4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc4 fma(float a, float b, acc4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int v) { return 4 * g + v; }
};

template <class T>
__global__ void __launch_bounds__(kThreads) k_gemm_update(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                         int mode, int p, int nb) {
  N = N.at(blockIdx.z);
  typedef typename MatrixCore<T>::acc4 acc4;
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
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = 32 * (wave & 1), wc = 32 * (wave >> 1);  // the wave's quarter of the tile
  const int l16 = lane & 15, lk = lane >> 4;
  acc4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = acc4{(T)0, (T)0, (T)0, (T)0};
  const int rbase = row0 + ti * kTile, cbase = col0 + tj * kTile;
  const int clim = mode == 0 ? ncol : m;
  // staging: thread (lr, lt) carries row / column lr of the tile for k offsets lt + 4 v; the next
  // K step's values are loaded into registers before the current step's products (addresses
  // clamped into the front, values masked when stored)
  const int lr = tid & 63, lt = tid >> 6;
  const int gr = rbase + lr, gc = cbase + lr;
  const bool rin = gr < m, cin = gc < clim;
  const T* Ar = A + min(gr, m - 1);
  const T* Br = B + min(gc, m - 1);
  T pa[kKStep / 4], pb[kKStep / 4];
  auto stage = [&](int kk) {
#pragma unroll
    for (int v = 0; v < kKStep / 4; ++v) {
      const int64_t gk = k0 + min(kk + lt + 4 * v, kn - 1);
      pa[v] = Ar[gk * m];
      pb[v] = Br[gk * m];
    }
  };
  stage(0);
  for (int kk = 0; kk < kn; kk += kKStep) {
#pragma unroll
    for (int v = 0; v < kKStep / 4; ++v) {
      const int t = lt + 4 * v;
      const bool kin = kk + t < kn;
      As[t][lr] = (kin && rin) ? pa[v] : (T)0;
      Bs[t][lr] = (kin && cin) ? pb[v] : (T)0;
    }
    __syncthreads();
    if (kk + kKStep < kn) stage(kk + kKStep);
#pragma unroll
    for (int q = 0; q < kKStep / 4; ++q) {
      const int kq = 4 * q + lk;
      T wv[2], lv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) wv[a] = As[kq][wr + 16 * a + l16];
#pragma unroll
      for (int b = 0; b < 2; ++b) lv[b] = Bs[kq][wc + 16 * b + l16];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = MatrixCore<T>::fma(lv[b], wv[a], acc[a][b]);
    }
    __syncthreads();
  }
  // C -= acc on the lower triangle: all 16 targets are read before the first is written. The
  // contribution block (mode 1) receives the children's entries only afterwards, so there the
  // product is stored as it is: C = -acc, with nothing read and no zero fill beforehand.
  const int cb = m - ncol;
  T* const Cbase = mode == 0 ? N.L + S.panel_off[s] : N.CB + S.cb_off[s];
  T* dst[2][2][4];
  T cv[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = rbase + wr + 16 * a + l16, c = cbase + wc + 16 * b + MatrixCore<T>::row(lk, v);
        const bool ok = r < m && r >= c && c < clim;
        dst[a][b][v] = ok ? Cbase + (mode == 0 ? (int64_t)c * m + r : (int64_t)(c - ncol) * cb + (r - ncol)) : nullptr;
        cv[a][b][v] = (ok && mode == 0) ? *dst[a][b][v] : (T)0;
      }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (dst[a][b][v]) *dst[a][b][v] = cv[a][b][v] - acc[a][b][v];
}

// ------------------------------------------------------------------------------ host engine

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

// Device blocks of the analysis and of factorization objects are kept for reuse instead of being
// returned to the driver: repeated solves of a pattern allocate the same sizes, and after a few
// hundred GB of hipMalloc / hipFree cycles the driver's calls took up to seconds each. A cached
// block is handed out again only for exactly its size. A block freed with the streams that used it
// carries one event per stream, recorded at the free: its next owner waits for those events (the
// work queued on those streams before the free), not for the whole device, so a worker that frees
// its factorization does not stall the others. A block freed without streams synchronizes the
// device first (as hipFree does). At most cap() idle bytes stay cached per device (the longest-idle
// blocks go back first; HSG_CACHE_CAP_GB sets the cap, default 32 GiB); hsg_release_cached_memory
// returns every idle block, and when hipMalloc runs out of memory the device's idle blocks are
// returned and the call is retried.
class BlockCache {
 public:
  void* alloc(size_t bytes, const char* what) {
    int dev = 0;
    hip_check(hipGetDevice(&dev), what);
    Block hit;
    bool found = false;
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (auto it = idle_.begin(); it != idle_.end(); ++it)
        if (it->dev == dev && it->bytes == bytes) {
          hit = *it;
          idle_bytes_[dev] -= bytes;
          idle_.erase(it);
          live_[hit.ptr] = {dev, bytes};
          found = true;
          break;
        }
    }
    if (found) {
      wait_events(hit);
      return hit.ptr;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      release_idle(dev);
      e = hipMalloc(&p, bytes);
    }
    hip_check(e, what);
    std::lock_guard<std::mutex> lk(mu_);
    live_[p] = {dev, bytes};
    return p;
  }
  // streams: the streams whose queued work may still use the block (empty: synchronize the device)
  void free(void* p, const std::vector<hipStream_t>& streams) {
    Block b;
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = live_.find(p);
      if (it == live_.end()) return;
      b.dev = it->second.dev;
      b.bytes = it->second.bytes;
      b.ptr = p;
      live_.erase(it);
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != b.dev) (void)hipSetDevice(b.dev);
    bool synced = streams.empty();
    for (hipStream_t st : streams) {
      hipEvent_t ev = nullptr;
      if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev, st) != hipSuccess) {
        if (ev) (void)hipEventDestroy(ev);
        (void)hipGetLastError();
        synced = true;  // no event: fall back to a device synchronization
        break;
      }
      b.events.push_back(ev);
    }
    if (synced) {
      for (hipEvent_t ev : b.events) (void)hipEventDestroy(ev);
      b.events.clear();
      (void)hipDeviceSynchronize();
    }
    std::vector<Block> evict;
    {
      std::lock_guard<std::mutex> lk(mu_);
      idle_.push_back(b);
      idle_bytes_[b.dev] += b.bytes;
      for (auto it = idle_.begin(); it != idle_.end() && idle_bytes_[b.dev] > cap();) {
        if (it->dev == b.dev) {
          idle_bytes_[b.dev] -= it->bytes;
          evict.push_back(*it);
          it = idle_.erase(it);
        } else {
          ++it;
        }
      }
    }
    for (Block& q : evict) {
      wait_events(q);
      (void)hipFree(q.ptr);
    }
    if (cur != b.dev) (void)hipSetDevice(cur);
  }
  // Returns the idle blocks of the current device to the driver; the bytes returned.
  size_t release_current() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    return release_idle(dev);
  }

 private:
  struct Block {
    int dev = 0;
    size_t bytes = 0;
    void* ptr = nullptr;
    std::vector<hipEvent_t> events;  // work that may still use the block
  };
  struct Owner {
    int dev;
    size_t bytes;
  };
  static size_t cap() {
    static const size_t c = [] {
      const char* e = std::getenv("HSG_CACHE_CAP_GB");
      const double gb = e ? std::atof(e) : 32.0;
      return gb > 0 ? (size_t)(gb * double(size_t(1) << 30)) : size_t(0);
    }();
    return c;
  }
  static void wait_events(Block& b) {
    for (hipEvent_t ev : b.events) {
      (void)hipEventSynchronize(ev);
      (void)hipEventDestroy(ev);
    }
    b.events.clear();
  }
  size_t release_idle(int dev) {
    std::vector<Block> out;
    size_t bytes = 0;
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (auto it = idle_.begin(); it != idle_.end();) {
        if (it->dev == dev) {
          bytes += it->bytes;
          out.push_back(*it);
          it = idle_.erase(it);
        } else {
          ++it;
        }
      }
      idle_bytes_[dev] = 0;
    }
    for (Block& q : out) {
      wait_events(q);
      (void)hipFree(q.ptr);
    }
    return bytes;
  }
  std::mutex mu_;
  std::unordered_map<void*, Owner> live_;
  std::list<Block> idle_;  // oldest first
  std::unordered_map<int, size_t> idle_bytes_;
};

BlockCache& block_cache() {
  static BlockCache* c = new BlockCache();  // never destroyed: frees may run during process exit
  return *c;
}

void* dalloc(size_t bytes, const char* what) { return block_cache().alloc(std::max<size_t>(1, bytes), what); }

template <class P>
void dalloc_into(P*& p, size_t bytes, const char* what) {
  p = static_cast<P*>(dalloc(bytes, what));
}

template <class V>
V* upload(const std::vector<V>& h) {
  const size_t bytes = std::max<size_t>(1, h.size()) * sizeof(V);
  V* d = static_cast<V*>(dalloc(bytes, "hipMalloc (analysis)"));
  if (!h.empty()) hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice), "hipMemcpy (analysis)");
  return d;
}

void dfree(void* p) {
  if (p) block_cache().free(p, {});
}

}  // namespace

size_t release_cached_blocks() { return block_cache().release_current(); }

namespace {

// frees a block that only work queued on `streams` may still use
void dfree_after(void* p, const std::vector<hipStream_t>& streams) {
  if (p) block_cache().free(p, streams);
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
  // W = L D workspace of the large fronts: used only while the front's level is factorized, so
  // the levels share one array (offsets restart at 0 on every level; its size is the largest
  // level's sum)
  h_w_off_.assign(nsn, -1);
  for (int l = 0; l < sym_.nlevels; ++l) {
    int64_t off = 0;
    for (int q = sym_.level_ptr[l]; q < sym_.level_ptr[l + 1]; ++q) {
      const int s = sym_.level_sn[q];
      if (sym_.sn_m[s] > opt_.small_front_max) {
        h_w_off_[s] = off;
        off += (int64_t)sym_.sn_m[s] * sym_.sn_ncol[s];
      }
    }
    w_total_ = std::max(w_total_, off);
  }
  panel_total_ = sym_.panel_off[nsn];
  cb_total_ = sym_.cb_total;
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
    for (int s2 : L.h_all) {
      L.max_m = std::max(L.max_m, sym_.sn_m[s2]);
      L.max_ncol = std::max(L.max_ncol, sym_.sn_ncol[s2]);
    }
    // algorithmic flops of the front-update products (lower-triangle outputs only): trailing
    // update after panel p: 2 kn sum_{c = c1}^{ncol - 1} (m - c); contribution block:
    // 2 ncol cb (cb + 1) / 2
    const int nbw = opt_.panel_width;
    L.upd_flops.assign((L.large_max_ncol + nbw - 1) / nbw, 0.0);
    for (int s2 : L.h_large) {
      const double m = sym_.sn_m[s2], nc = sym_.sn_ncol[s2];
      for (int pp = 0; pp < (int)L.upd_flops.size(); ++pp) {
        const double c0 = (double)pp * nbw, c1 = std::min(nc, c0 + nbw);
        if (c0 >= nc || c1 >= nc) continue;
        const double cols = nc - c1;
        const double rows_sum = cols * m - 0.5 * (c1 + nc - 1) * cols;  // sum_{c=c1}^{nc-1} (m - c)
        L.upd_flops[pp] += 2.0 * (c1 - c0) * rows_sum;
      }
      const double cbn = m - nc;
      L.cb_flops += nc * cbn * (cbn + 1.0);
    }
    for (double f : L.upd_flops) update_flops_ += f;
    update_flops_ += L.cb_flops;
    L.n_all = (int)L.h_all.size();
    L.n_small = (int)L.h_small.size();
    L.n_large = (int)L.h_large.size();
    L.fwd_tiles = (L.max_m + kGvTile - 1) / kGvTile;
    // columns per chunk of the split products: on levels with fronts of more than 2 kGvSplit
    // pivots, halved from kGvSplit (down to 64) while the level's grid stays below ~6 workgroups
    // per CU, so that the few large fronts of the top levels spread their G reads over the whole
    // device; smaller fronts keep kGvSplit (a split costs a second launch, which small problems,
    // bound by launches, cannot afford)
    L.fwd_split = kGvSplit;
    while (L.max_ncol > 2 * kGvSplit && L.fwd_split > 64 &&
           (int64_t)L.fwd_tiles * L.n_all * ((L.max_ncol + L.fwd_split - 1) / L.fwd_split) < kFwdTargetGroups)
      L.fwd_split /= 2;
    L.fwd_chunks = (L.max_ncol + L.fwd_split - 1) / L.fwd_split;
    n_small_ += L.n_small;
    n_large_ += L.n_large;
    max_large_per_level_ = std::max(max_large_per_level_, L.n_large);
  }
  plan_solve();
}

void Engine::plan_solve() {
  const int nsn = sym_.nsn;
  // diagonal 64-blocks of every L11
  for (int s = 0; s < nsn; ++s) {
    const int nb2 = (sym_.sn_ncol[s] + kSolveBlock - 1) / kSolveBlock;
    for (int b = 0; b < nb2; ++b) {
      h_pairs_.push_back(s);
      h_pairs_.push_back(b);
    }
  }
  n_pairs_ = (int)(h_pairs_.size() / 2);
  for (int s = 0; s < nsn; ++s) {
    const double nc = sym_.sn_ncol[s], cb = sym_.sn_m[s] - sym_.sn_ncol[s];
    g_flops_ += cb * nc * nc + nc * nc * nc / 3.0;
  }
  auto task = [](int64_t a, int32_t lda, int8_t abuf, int64_t b, int32_t ldb, int8_t bbuf, int64_t c, int32_t ldc,
                 int8_t cbuf, int rows, int cols, int k0, int k1, bool neg) {
    GemmTask t;
    t.a = a;
    t.b = b;
    t.c = c;
    t.lda = lda;
    t.ldb = ldb;
    t.ldc = ldc;
    t.rows = rows;
    t.cols = cols;
    t.k0 = k0;
    t.k1 = k1;
    t.abuf = abuf;
    t.bbuf = bbuf;
    t.cbuf = cbuf;
    t.neg = neg ? 1 : 0;
    return t;
  };
  constexpr int8_t kBufL = 0, kBufG = 1, kBufT = 2;
  // recursive doubling: at half size h, every pair A = [A0, A0 + h), B = [A0 + h, A0 + h + nB):
  //   T = L(B, A) G(A, A)      (scratch, nB x h, ld nB)
  //   G(B, A) = -G(B, B) T
  for (int64_t h = kSolveBlock; h < sym_.max_ncol; h *= 2) {
    std::vector<GemmTask> b1, b2;
    int64_t toff = 0;
    for (int s = 0; s < nsn; ++s) {
      const int ncol = sym_.sn_ncol[s], m = sym_.sn_m[s];
      const int64_t po = sym_.panel_off[s];
      for (int64_t A0 = 0; A0 + h < ncol; A0 += 2 * h) {
        const int64_t B0 = A0 + h;
        const int nB = (int)std::min<int64_t>(h, ncol - B0);
        for (int ri = 0; ri * 64 < nB; ++ri)
          for (int ci = 0; ci * 64 < h; ++ci) {
            const int rows = std::min(64, nB - ri * 64);
            b1.push_back(task(po + A0 * m + B0 + ri * 64, m, kBufL, po + (A0 + ci * 64) * m + A0, m, kBufG,
                              toff + (int64_t)ci * 64 * nB + ri * 64, nB, kBufT, rows, 64, ci * 64, (int)h, false));
            b2.push_back(task(po + B0 * m + B0 + ri * 64, m, kBufG, toff + (int64_t)ci * 64 * nB, nB, kBufT,
                              po + (A0 + ci * 64) * m + B0 + ri * 64, m, kBufG, rows, 64, 0,
                              std::min(nB, ri * 64 + 64), true));
          }
        toff += (int64_t)nB * h;
      }
    }
    t_total_ = std::max(t_total_, toff);
    for (auto* v : {&b1, &b2}) {
      gstages_.push_back({(int64_t)h_tasks_.size(), (int64_t)v->size()});
      h_tasks_.insert(h_tasks_.end(), v->begin(), v->end());
    }
  }
  // G21 = -L21 G11
  {
    const int64_t begin = (int64_t)h_tasks_.size();
    for (int s = 0; s < nsn; ++s) {
      const int ncol = sym_.sn_ncol[s], m = sym_.sn_m[s], cb = m - ncol;
      const int64_t po = sym_.panel_off[s];
      for (int ri = 0; ri * 64 < cb; ++ri)
        for (int ci = 0; ci * 64 < ncol; ++ci)
          h_tasks_.push_back(task(po + ncol + ri * 64, m, kBufL, po + (int64_t)ci * 64 * m, m, kBufG,
                                  po + (int64_t)ci * 64 * m + ncol + ri * 64, m, kBufG, std::min(64, cb - ri * 64),
                                  std::min(64, ncol - ci * 64), ci * 64, ncol, true));
    }
    gstages_.push_back({begin, (int64_t)h_tasks_.size() - begin});
  }
  // row-major copy: tiles on and below the diagonal tile
  for (int s = 0; s < nsn; ++s) {
    const int ncol = sym_.sn_ncol[s], m = sym_.sn_m[s];
    for (int ti = 0; ti * 64 < m; ++ti)
      for (int tj = 0; tj * 64 < ncol && tj <= ti; ++tj) {
        h_ttiles_.push_back(s);
        h_ttiles_.push_back(ti);
        h_ttiles_.push_back(tj);
      }
  }
  n_ttiles_ = (int)(h_ttiles_.size() / 3);
  // forward tasks per level: 64-row tiles of every front, each tile's columns (up to the tile's
  // end inside G11) in chunks of the level's split width; split tiles store partial sums in slots
  // numbered per level
  for (Level& L : levels_) {
    L.ft_begin = (int64_t)h_ftasks_.size();
    L.fh_begin = (int64_t)h_fheads_.size();
    int64_t slots = 0;
    for (int s : L.h_all) {
      const int ncol = sym_.sn_ncol[s], m = sym_.sn_m[s];
      for (int r0 = 0; r0 < m; r0 += kGvTile) {
        const int kend = std::min(ncol, r0 + kGvTile);
        const int nchunk = (kend + L.fwd_split - 1) / L.fwd_split;
        FwdTask t{};
        t.m = m;
        t.ncol = ncol;
        t.first = sym_.sn_first[s];
        t.r0 = r0;
        t.nchunk = nchunk;
        t.goff = sym_.panel_off[s];
        t.rowoff = sym_.sn_rowoff[s];
        for (int c = 0; c < nchunk; ++c) {
          t.kbeg = c * L.fwd_split;
          t.kstop = std::min(kend, (c + 1) * L.fwd_split);
          t.slot = nchunk > 1 ? slots + c : -1;
          h_ftasks_.push_back(t);
          if (c == 0 && nchunk > 1) h_fheads_.push_back(t);
        }
        if (nchunk > 1) slots += nchunk;
      }
    }
    L.ft_count = (int64_t)h_ftasks_.size() - L.ft_begin;
    L.fh_count = (int64_t)h_fheads_.size() - L.fh_begin;
    // the launch's register tile: columns per thread and pass, rounded up to a power of two
    int cols = 1;
    for (int64_t k = L.ft_begin; k < L.ft_begin + L.ft_count; ++k)
      cols = std::max(cols, (std::min(kGvChunk, h_ftasks_[k].kstop - h_ftasks_[k].kbeg) + 3) / 4);
    L.ft_cols = 2;
    while (L.ft_cols < cols) L.ft_cols *= 2;
    // levels whose tiles all have at most kWaveCols columns in one chunk: one task per wave, a
    // lane per row, C = the tiles' columns (rounded up to a power of two >= 8)
    L.ft_wave = L.ft_count > 0;
    int wcols = 1;
    for (int64_t k = L.ft_begin; k < L.ft_begin + L.ft_count; ++k) {
      const FwdTask& t = h_ftasks_[k];
      if (t.nchunk != 1 || t.kstop - t.kbeg > kWaveCols) L.ft_wave = false;
      wcols = std::max(wcols, t.kstop - t.kbeg);
    }
    if (L.ft_wave) {
      L.ft_cols = 8;
      while (L.ft_cols < wcols) L.ft_cols *= 2;
    }
    part_slots_ = std::max(part_slots_, slots);
  }
  // backward tasks per level: column tiles of at most 64 (width rounded up to a power of two >= 8:
  // the thread layout), rows from the tile's first column to m in chunks of 32 rows per row group
  // (at most kBwdRows rows: the staging area). Chunks of 16 rows (shorter load chains per task,
  // more partial sums) were measured slower once the partials were added by the next level's
  // tasks (889 against 880 us per box48 solve).
  auto tile_width = [](int jn) {
    int jw = 8;
    while (jw < jn) jw *= 2;
    return jw;
  };
  for (Level& L : levels_) {
    L.bt_begin = (int64_t)h_btasks_.size();
    L.bh_begin = (int64_t)h_bheads_.size();
    int64_t slots = 0;
    L.bt_rows = 32;
    for (int s : L.h_all) {
      const int ncol = sym_.sn_ncol[s], m = sym_.sn_m[s];
      for (int j0 = 0; j0 < ncol; j0 += kGvTile) {
        const int jn = std::min(kGvTile, ncol - j0);
        const int jw = tile_width(jn);
        const int rc = std::min(kBwdRows, L.bt_rows * (256 / jw));
        const int rows = m - j0;
        const int nchunk = (rows + rc - 1) / rc;
        BwdTask t{};
        t.s = s;
        t.j0 = j0;
        t.jn = jn;
        t.jw = jw;
        t.nchunk = nchunk;
        t.slot0 = nchunk > 1 ? slots : -1;
        t.ncol = ncol;
        t.first = sym_.sn_first[s];
        for (int c = 0; c < nchunk; ++c) {
          t.chunk = c;
          t.i0 = j0 + c * rc;
          t.in = std::min(rc, m - t.i0);
          t.goff = sym_.panel_off[s] + (int64_t)t.i0 * ncol + j0;
          t.rpoff = sym_.sn_rowoff[s] + t.i0;
          h_btasks_.push_back(t);
          if (c == 0 && nchunk > 1) h_bheads_.push_back(t);
        }
        if (nchunk > 1) slots += nchunk;
      }
    }
    L.bt_count = (int64_t)h_btasks_.size() - L.bt_begin;
    L.bh_count = (int64_t)h_bheads_.size() - L.bh_begin;
    int rows = 1;
    for (int64_t k = L.bt_begin; k < L.bt_begin + L.bt_count; ++k)
      rows = std::max(rows, (h_btasks_[k].in + 256 / h_btasks_[k].jw - 1) / (256 / h_btasks_[k].jw));
    L.bt_regs = 4;
    while (L.bt_regs < rows) L.bt_regs *= 2;
    part_slots_ = std::max(part_slots_, slots);
  }
  // per pivot position: the split tile holding it between its level's launch and the next one
  h_psplit_.assign(sym_.n, PivSplit{-1, 0, 0, -1});
  for (int l = 0; l < (int)levels_.size(); ++l) {
    const Level& L = levels_[l];
    for (int s2 : L.h_all)
      for (int j = 0; j < sym_.sn_ncol[s2]; ++j) h_psplit_[sym_.sn_first[s2] + j].level = l;
    for (int64_t k = L.bh_begin; k < L.bh_begin + L.bh_count; ++k) {
      const BwdTask& h = h_bheads_[k];
      for (int j2 = 0; j2 < h.jn; ++j2) {
        PivSplit& q = h_psplit_[h.first + h.j0 + j2];
        q.slot0 = (int32_t)h.slot0;
        q.nchunk = h.nchunk;
        q.jj = j2;
      }
    }
  }
  // pull map of the forward products: for front row r (index rowoff + r), the children's U rows
  // mapped onto it, children in child-list order
  h_cptr_.assign(rows_total_ + 1, 0);
  for (int s = 0; s < nsn; ++s)
    for (int q = sym_.child_ptr[s]; q < sym_.child_ptr[s + 1]; ++q) {
      const int ch = sym_.child_list[q];
      const int64_t coff = sym_.sn_rowoff[ch] + sym_.sn_ncol[ch];
      const int cbc = sym_.sn_m[ch] - sym_.sn_ncol[ch];
      for (int r = 0; r < cbc; ++r) ++h_cptr_[sym_.sn_rowoff[s] + sym_.relmap[coff + r] + 1];
    }
  for (int64_t r = 0; r < rows_total_; ++r) h_cptr_[r + 1] += h_cptr_[r];
  h_cidx_.assign(std::max<int64_t>(1, h_cptr_[rows_total_]), 0);
  {
    std::vector<int64_t> cur(h_cptr_.begin(), h_cptr_.end() - 1);
    for (int s = 0; s < nsn; ++s)
      for (int q = sym_.child_ptr[s]; q < sym_.child_ptr[s + 1]; ++q) {
        const int ch = sym_.child_list[q];
        const int64_t coff = sym_.sn_rowoff[ch] + sym_.sn_ncol[ch];
        const int cbc = sym_.sn_m[ch] - sym_.sn_ncol[ch];
        for (int r = 0; r < cbc; ++r) h_cidx_[cur[sym_.sn_rowoff[s] + sym_.relmap[coff + r]]++] = (int32_t)(coff + r);
      }
  }
  if (rows_total_ >= INT32_MAX) throw std::runtime_error("front row lists too long for 32-bit U indices");
  // the same map with the first two children's rows inline (most rows have at most two), so that a
  // gather reads them without the indirection through cidx
  h_cpull_.assign(rows_total_, PullRow{-1, -1, 0, 0});
  for (int64_t r = 0; r < rows_total_; ++r) {
    const int64_t b = h_cptr_[r], e = h_cptr_[r + 1];
    PullRow& q = h_cpull_[r];
    if (e > b) q.u0 = h_cidx_[b];
    if (e > b + 1) q.u1 = h_cidx_[b + 1];
    q.b = (int32_t)std::min(e, b + 2);
    q.e = (int32_t)e;
  }
  validate_solve_plan();
}

// Every element a planned launch of the solve phase may touch lies inside its buffer: tile
// products against the panel and scratch sizes, split products against their scratch. Throws
// (no launch happens) on a violation.
void Engine::validate_solve_plan() const {
  auto bad = [](const char* what) { throw std::runtime_error(std::string("solve plan: ") + what); };
  const int64_t sizes[3] = {panel_total_, panel_total_, t_total_};
  for (const GemmTask& t : h_tasks_) {
    if (t.abuf < 0 || t.abuf > 2 || t.bbuf < 0 || t.bbuf > 2 || t.cbuf < 1 || t.cbuf > 2) bad("buffer id");
    if (t.rows < 1 || t.rows > kGmTile || t.cols < 1 || t.cols > kGmTile || t.k0 < 0 || t.k1 <= t.k0) bad("tile shape");
    const int64_t a_lo = t.a + (int64_t)t.k0 * t.lda, a_hi = t.a + (int64_t)(t.k1 - 1) * t.lda + t.rows - 1;
    const int64_t b_lo = t.b + t.k0, b_hi = t.b + (int64_t)(t.cols - 1) * t.ldb + t.k1 - 1;
    const int64_t c_lo = t.c, c_hi = t.c + (int64_t)(t.cols - 1) * t.ldc + t.rows - 1;
    if (a_lo < 0 || a_hi >= sizes[t.abuf]) bad("A operand out of range");
    if (b_lo < 0 || b_hi >= sizes[t.bbuf]) bad("B operand out of range");
    if (c_lo < 0 || c_hi >= sizes[t.cbuf]) bad("C operand out of range");
  }
  for (const TaskRange& r : gstages_)
    if (r.begin < 0 || r.count < 0 || r.begin + r.count > (int64_t)h_tasks_.size()) bad("stage range");
  for (const Level& L : levels_) {
    if (L.n_all != (int)L.h_all.size()) bad("level size");
    if (L.ft_begin < 0 || L.ft_begin + L.ft_count > (int64_t)h_ftasks_.size()) bad("forward task range");
    if (L.fh_begin < 0 || L.fh_begin + L.fh_count > (int64_t)h_fheads_.size()) bad("forward split range");
    if (L.bt_regs != 4 && L.bt_regs != 8 && L.bt_regs != 16 && L.bt_regs != 32) bad("backward rows per thread");
    for (int64_t k = L.bt_begin; k < L.bt_begin + L.bt_count; ++k) {
      const BwdTask& t = h_btasks_[k];
      if (t.in > L.bt_regs * (256 / t.jw)) bad("backward task rows per thread");
    }
    if (L.ft_cols != 2 && L.ft_cols != 4 && L.ft_cols != 8 && L.ft_cols != 16 && L.ft_cols != 32)
      bad("forward columns per thread");
    for (int64_t k = L.ft_begin; k < L.ft_begin + L.ft_count; ++k) {
      const FwdTask& t = h_ftasks_[k];
      if (L.ft_wave ? (t.nchunk != 1 || t.kstop - t.kbeg > L.ft_cols || L.ft_cols < 8)
                    : (std::min(kGvChunk, t.kstop - t.kbeg) + 3) / 4 > L.ft_cols)
        bad("forward task columns per thread");
    }
    if (L.bt_begin < 0 || L.bt_begin + L.bt_count > (int64_t)h_btasks_.size()) bad("backward task range");
    if (L.bh_begin < 0 || L.bh_begin + L.bh_count > (int64_t)h_bheads_.size()) bad("backward split range");
  }
  if ((int64_t)h_cpull_.size() != rows_total_) bad("pull map size");
  for (const PullRow& q : h_cpull_)
    if (q.u0 >= rows_total_ || q.u1 >= rows_total_ || (q.u0 < 0 && q.u1 >= 0) || q.b < 0 || q.b > q.e ||
        q.e > (int64_t)h_cidx_.size())
      bad("pull map entry");
  for (const FwdTask& t : h_ftasks_) {
    if (t.m < 1 || t.ncol < 1 || t.ncol > t.m || t.first < 0 || t.first + t.ncol > sym_.n) bad("forward task front");
    if (t.r0 < 0 || t.r0 >= t.m || t.kbeg < 0 || t.kstop <= t.kbeg || t.kstop > std::min(t.ncol, t.r0 + kGvTile))
      bad("forward task range");
    if (t.goff < 0 || t.goff + (int64_t)t.m * t.ncol > panel_total_) bad("forward task panel");
    if (t.rowoff < 0 || t.rowoff + t.m > rows_total_) bad("forward task rows");
    if (t.nchunk > 1 ? (t.slot < 0 || t.slot >= part_slots_) : t.slot != -1) bad("forward task scratch");
  }
  for (const FwdTask& t : h_fheads_)
    if (t.nchunk < 2 || t.kbeg != 0 || t.slot < 0 || t.slot + t.nchunk > part_slots_) bad("forward split head");
  for (const BwdTask& t : h_btasks_) {
    if (t.s < 0 || t.s >= sym_.nsn) bad("backward task front");
    const int ncol = sym_.sn_ncol[t.s], m = sym_.sn_m[t.s];
    if (t.ncol != ncol || t.first != sym_.sn_first[t.s] || t.in > kBwdRows) bad("backward task front data");
    if (t.goff != sym_.panel_off[t.s] + (int64_t)t.i0 * ncol + t.j0 || t.rpoff != sym_.sn_rowoff[t.s] + t.i0)
      bad("backward task offsets");
    if (t.jw < 8 || t.jw > kGvTile || (t.jw & (t.jw - 1)) || t.jn < 1 || t.jn > t.jw) bad("backward task width");
    if (t.j0 < 0 || t.j0 + t.jn > ncol || t.i0 < t.j0 || t.in < 1 || t.i0 + t.in > m) bad("backward task range");
    if (t.nchunk > 1 && (t.slot0 < 0 || t.slot0 + t.nchunk > part_slots_)) bad("backward task scratch");
  }
  for (const BwdTask& t : h_bheads_)
    if (t.chunk != 0 || t.nchunk < 2 || t.slot0 < 0 || t.slot0 + t.nchunk > part_slots_) bad("backward split head");
  if ((int64_t)h_psplit_.size() != sym_.n) bad("split state size");
  for (const PivSplit& q : h_psplit_)
    if (q.level < 0 || q.level >= (int)levels_.size() || q.jj < 0 || q.jj >= kGvTile ||
        (q.slot0 >= 0 && (q.nchunk < 2 || q.slot0 + q.nchunk > part_slots_)))
      bad("split state entry");
}

void Engine::ensure_device() {
  std::lock_guard<std::mutex> lk(device_mu_);
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
  d_pairs_ = upload(h_pairs_);
  d_ttiles_ = upload(h_ttiles_);
  d_tasks_ = upload(h_tasks_);
  d_ftasks_ = upload(h_ftasks_);
  d_fheads_ = upload(h_fheads_);
  d_btasks_ = upload(h_btasks_);
  d_bheads_ = upload(h_bheads_);
  d_psplit_ = upload(h_psplit_);
  d_cpull_ = upload(h_cpull_);
  d_cidx_ = upload(h_cidx_);
  for (Level& L : levels_) {
    L.d_all = upload(L.h_all);
    if (L.n_small) L.d_small = upload(L.h_small);
    if (L.n_large) L.d_large = upload(L.h_large);
  }
  on_device_ = true;
}

Engine::~Engine() {
  for (void* p : {(void*)d_first_, (void*)d_ncol_, (void*)d_m_, (void*)d_child_ptr_, (void*)d_child_list_,
                  (void*)d_relmap_, (void*)d_rows_orig_, (void*)d_rowoff_, (void*)d_panel_off_, (void*)d_cb_off_,
                  (void*)d_w_off_, (void*)d_asm_, (void*)d_ia_, (void*)d_ib_, (void*)d_pairs_, (void*)d_ttiles_,
                  (void*)d_tasks_, (void*)d_ftasks_, (void*)d_fheads_, (void*)d_btasks_, (void*)d_bheads_, (void*)d_psplit_, (void*)d_cpull_,
                  (void*)d_cidx_})
    dfree(p);
  for (auto& L : levels_) {
    dfree(L.d_all);
    dfree(L.d_small);
    dfree(L.d_large);
  }
}

Numeric::~Numeric() { free_all(); }

void Numeric::note_stream(hipStream_t st) {
  for (hipStream_t s : streams_)
    if (s == st) return;
  streams_.push_back(st);
}

void Numeric::free_all() {
  for (void* p : {d_L_, d_CB_, d_W_, d_dval_, d_doff_, (void*)d_ptype_, (void*)d_piv_, (void*)d_ipiv_,
                  (void*)d_bperm_, (void*)d_cnt_, d_Xp_, d_Y_, d_U_, d_G_, d_Gt_, d_T_, d_part_,
                  (void*)d_pos_, (void*)d_rpos_, (void*)d_zpull_})
    dfree_after(p, streams_);
  streams_.clear();
  d_L_ = d_CB_ = d_W_ = d_dval_ = d_doff_ = d_Xp_ = d_Y_ = d_U_ = d_G_ = d_Gt_ = d_T_ = d_part_ = nullptr;
  d_pos_ = d_rpos_ = nullptr;
  d_zpull_ = nullptr;
  g_ready_ = false;
  d_ptype_ = nullptr;
  d_piv_ = d_ipiv_ = d_bperm_ = nullptr;
  d_cnt_ = nullptr;
  ws_cols_ = 0;
  precision_ = -1;
  valid_ = false;
}

void Engine::release(Numeric& nm) { nm.free_all(); }

void Engine::alloc_numeric(Numeric& nm, int precision, int batch) {
  if (nm.precision_ == precision && nm.batch_ == batch && nm.d_L_) return;
  nm.free_all();
  const size_t ts = precision == 0 ? sizeof(double) : sizeof(float);
  const int64_t n = sym_.n;
  const size_t z = (size_t)std::max(1, batch);
  dalloc_into(nm.d_L_, z * std::max<int64_t>(1, panel_total_) * ts, "hipMalloc (factor panels)");
  dalloc_into(nm.d_CB_, z * std::max<int64_t>(1, cb_total_) * ts, "hipMalloc (contribution blocks)");
  dalloc_into(nm.d_W_, z * std::max<int64_t>(1, w_total_) * ts, "hipMalloc (panel workspace)");
  dalloc_into(nm.d_dval_, z * n * ts, "hipMalloc (D)");
  dalloc_into(nm.d_doff_, z * n * ts, "hipMalloc (D)");
  dalloc_into(nm.d_ptype_, z * n, "hipMalloc (pivots)");
  dalloc_into(nm.d_piv_, z * n * sizeof(int32_t), "hipMalloc (pivots)");
  dalloc_into(nm.d_ipiv_, z * n * sizeof(int32_t), "hipMalloc (pivots)");
  dalloc_into(nm.d_bperm_, z * bperm_stride() * sizeof(int32_t), "hipMalloc (pivots)");
  dalloc_into(nm.d_cnt_, z * kCntSlots * sizeof(unsigned long long), "hipMalloc (counters)");
  nm.precision_ = precision;
  nm.batch_ = (int)z;
}

// a kernel launch counted in `nl` (launches of a timed phase)
#define HSG_LAUNCH_COUNTED(...)     \
  do {                              \
    hipLaunchKernelGGL(__VA_ARGS__); \
    ++nl;                           \
  } while (0)

template <class T>
void Engine::factorize_t(Numeric& nm, int Z, const BatchValues& bv, double null_tol, hipStream_t st,
                         InertiaResult* out) {
  DevSym S{sym_.n, sym_.nsn, d_first_, d_ncol_, d_m_, d_rowoff_, d_panel_off_, d_cb_off_, d_w_off_,
           d_child_ptr_, d_child_list_, d_relmap_, d_rows_orig_};
  DevNum<T> N{(T*)nm.d_L_, (T*)nm.d_CB_, (T*)nm.d_W_, (T*)nm.d_dval_, (T*)nm.d_doff_, nm.d_ptype_, nm.d_piv_, nm.d_ipiv_, nm.d_cnt_};
  N.bperm = nm.d_bperm_;
  N.zL = std::max<int64_t>(1, panel_total_);
  N.zCB = std::max<int64_t>(1, cb_total_);
  N.zW = std::max<int64_t>(1, w_total_);
  N.zn = sym_.n;
  N.zbp = bperm_stride();
  const int nb = opt_.panel_width;
  const int detect = null_tol > 0 ? 1 : 0;
  const int64_t tok = timing_begin(kPhaseFactor, st);
  int64_t nl = 0;
  const int64_t lz = N.zL * Z;
  HSG_LAUNCH_COUNTED(k_zero<T>, dim3(blocks_for(lz / (16 / sizeof(T)) + 1)), dim3(256), 0, st, N.L, lz);
  HSG_LAUNCH_COUNTED(k_fill_u64, dim3((Z * kCntSlots + 63) / 64), dim3(64), 0, st, nm.d_cnt_, Z * (int)kCntSlots);
  if (nnz_ > 0)
    HSG_LAUNCH_COUNTED(k_assemble_original<T>, dim3(blocks_for(nnz_, 256, 1024), Z), dim3(kThreads), 0, st, nnz_, d_asm_,
                       d_ia_, d_ib_, bv, N.L, N.zL, nm.d_cnt_);
  for (const Level& L : levels_) {
    if (L.n_small) {
      const int m = L.small_maxm;
      const size_t shm = (size_t)m * m * sizeof(T) + 4 * (size_t)m * sizeof(T) + (size_t)m * sizeof(int) + m + 16;
      HSG_LAUNCH_COUNTED(k_factor_small<T>, dim3(L.n_small, 1, Z), dim3(kThreads), shm, st, S, N, L.d_small, null_tol,
                         detect);
    }
    if (L.n_large) {
      HSG_LAUNCH_COUNTED(k_extend_add_large<T>, dim3((L.large_max_ncol + kEaCols - 1) / kEaCols, L.n_large, Z),
                         dim3(kThreads), 0, st, S, N, L.d_large, 0);
      const int npan = (L.large_max_ncol + nb - 1) / nb;
      const size_t shm_bk = (size_t)nb * nb * sizeof(T) + 4 * (size_t)nb * sizeof(T) + nb * sizeof(int) + nb + 16;
      const size_t shm_pa = ((size_t)64 * 4 * kLqLd + (size_t)kPanelRows * 64) * sizeof(T);
      const int nrc = (L.large_maxm + kPanelRows - 1) / kPanelRows;
      const int cpg = std::max(1, kThreads / nb);
      const int nsw = (L.large_max_ncol + cpg - 1) / cpg;
      for (int p = 0; p < npan; ++p) {
        if constexpr (std::is_same<T, double>::value)
          HSG_LAUNCH_COUNTED(k_bk_wave, dim3(L.n_large, 1, Z), dim3(64 * kBkWaveLoaders), 0, st, S, N, L.d_large, p,
                             nb, null_tol, detect);
        else
          HSG_LAUNCH_COUNTED(k_bk_diag<T>, dim3(L.n_large, 1, Z), dim3(64 * kBkWaves), 0, st, S, N, L.d_large, p, nb,
                             null_tol, detect);
        HSG_LAUNCH_COUNTED(k_panel_apply<T>, dim3(nrc + nsw, L.n_large, Z), dim3(kThreads), shm_pa, st, S, N, L.d_large,
                           p, nb);
        const int rem = L.large_maxm - (p + 1) * nb;
        if (rem > 0 && (p + 1) * nb < L.large_max_ncol) {
          const int t = (rem + kTile - 1) / kTile;
          const int tc = (L.large_max_ncol - (p + 1) * nb + kTile - 1) / kTile;
          const int64_t ut = timing_begin(kPhaseUpdate, st);
          HSG_LAUNCH_COUNTED(k_gemm_update<T>, dim3(t * tc, L.n_large, Z), dim3(kThreads), 0, st, S, N, L.d_large, 0,
                             p, nb);
          timing_end(ut, st, 1, L.upd_flops[p]);
        }
      }
      if (L.large_max_cb > 0) {
        const int t = (L.large_max_cb + kTile - 1) / kTile;
        const int64_t ut = timing_begin(kPhaseUpdate, st);
        HSG_LAUNCH_COUNTED(k_gemm_update<T>, dim3(t * t, L.n_large, Z), dim3(kThreads), 0, st, S, N, L.d_large, 1, 0,
                           nb);
        timing_end(ut, st, 1, L.cb_flops);
        HSG_LAUNCH_COUNTED(k_extend_add_large<T>, dim3((L.large_max_cb + kEaCols - 1) / kEaCols, L.n_large, Z),
                           dim3(kThreads), 0, st, S, N, L.d_large, 1);
      }
    }
  }
  hip_check(hipGetLastError(), "factorization kernels");
  timing_end(tok, st, nl, sym_.flops * Z);
  std::vector<unsigned long long> h((size_t)Z * kCntSlots);
  hip_check(hipMemcpyAsync(h.data(), nm.d_cnt_, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
            "hipMemcpyAsync (counters)");
  hip_check(hipStreamSynchronize(st), "factorization");
  for (int z = 0; z < Z; ++z) {
    const unsigned long long* c = h.data() + (size_t)z * kCntSlots;
    InertiaResult& r = out[z];
    r.neg = (int64_t)c[kCntNeg];
    r.zero = (int64_t)c[kCntZero];
    r.pos = (int64_t)c[kCntPos];
    r.two = (int64_t)c[kCnt2x2];
    r.exact_zero = (int64_t)c[kCntExactZero];
    r.singular = r.zero > 0;
    if (r.neg + r.zero + r.pos != sym_.n) throw std::runtime_error("factorization: pivot counts do not add up to n");
  }
}

InertiaResult Engine::factorize(Numeric& nm, double alpha, double beta, const double* av, const double* bv, int precision,
                                double null_tol, hipStream_t stream) {
  InertiaResult r;
  factorize_batch(nm, 1, &alpha, &beta, &av, &bv, precision, null_tol, stream, &r);
  nm.valid_ = !r.singular;
  return r;
}

void Engine::factorize_batch(Numeric& nm, int Z, const double* alphas, const double* betas, const double* const* avs,
                             const double* const* bvs, int precision, double null_tol, hipStream_t stream,
                             InertiaResult* out) {
  if (Z < 1 || Z > kMaxBatch) throw std::runtime_error("factorize_batch: between 1 and 64 factorizations per batch");
  nm.valid_ = false;
  nm.g_ready_ = false;
  ensure_device();
  alloc_numeric(nm, precision, Z);
  nm.note_stream(stream);
  BatchValues bv{};
  for (int z = 0; z < Z; ++z) {
    bv.a[z] = avs[z];
    bv.b[z] = bvs[z];
    bv.alpha[z] = alphas[z];
    bv.beta[z] = betas[z];
  }
  if (precision == 0) factorize_t<double>(nm, Z, bv, null_tol, stream, out);
  else factorize_t<float>(nm, Z, bv, null_tol, stream, out);
}

template <class T>
void Engine::build_inverse_panels(Numeric& nm, hipStream_t st) {
  const size_t ts = sizeof(T);
  if (!nm.d_G_) dalloc_into(nm.d_G_, std::max<int64_t>(1, panel_total_) * ts, "hipMalloc (inverse panels)");
  if (!nm.d_Gt_) dalloc_into(nm.d_Gt_, std::max<int64_t>(1, panel_total_) * ts, "hipMalloc (inverse panels)");
  if (!nm.d_T_) dalloc_into(nm.d_T_, std::max<int64_t>(1, t_total_) * ts, "hipMalloc (inverse scratch)");
  if (!nm.d_pos_) dalloc_into(nm.d_pos_, (size_t)sym_.n * sizeof(int32_t), "hipMalloc (pivot positions)");
  if (!nm.d_zpull_)
    dalloc_into(nm.d_zpull_, (size_t)sym_.n * sizeof(PullRow), "hipMalloc (pivot positions)");
  if (!nm.d_rpos_)
    dalloc_into(nm.d_rpos_, (size_t)std::max<int64_t>(1, rows_total_) * sizeof(int32_t),
                "hipMalloc (pivot positions)");
  DevSym S{sym_.n, sym_.nsn, d_first_, d_ncol_, d_m_, d_rowoff_, d_panel_off_, d_cb_off_, d_w_off_,
           d_child_ptr_, d_child_list_, d_relmap_, d_rows_orig_};
  DevNum<T> N{(T*)nm.d_L_, (T*)nm.d_CB_, (T*)nm.d_W_, (T*)nm.d_dval_, (T*)nm.d_doff_, nm.d_ptype_, nm.d_piv_, nm.d_ipiv_, nm.d_cnt_};
  T* G = (T*)nm.d_G_;
  const int64_t tok = timing_begin(kPhaseInverse, st);
  int64_t nl = 0;
  if (n_pairs_ > 0) {
    hipLaunchKernelGGL(k_diag_block_inverse<T>, dim3(n_pairs_), dim3(64), 0, st, S, N, d_pairs_, G);
    ++nl;
  }
  for (const TaskRange& r : gstages_)
    if (r.count > 0) {
      hipLaunchKernelGGL(k_tile_gemm<T>, dim3((unsigned)r.count), dim3(256), 0, st, d_tasks_ + r.begin, N.L, G,
                         (T*)nm.d_T_);
      ++nl;
    }
  if (n_ttiles_ > 0) {
    hipLaunchKernelGGL(k_transpose_g<T>, dim3(n_ttiles_), dim3(256), 0, st, S, d_ttiles_, (const T*)G, (T*)nm.d_Gt_);
    ++nl;
  }
  // pivot positions of the rows after this factorization's local pivoting
  hipLaunchKernelGGL(k_pivot_positions<T>, dim3(sym_.nsn), dim3(256), 0, st, S, N, nm.d_pos_,
                     d_cpull_, nm.d_zpull_);
  hipLaunchKernelGGL(k_row_positions, dim3(blocks_for(rows_total_)), dim3(256), 0, st, rows_total_,
                     (const int32_t*)d_rows_orig_, (const int32_t*)nm.d_pos_, nm.d_rpos_);
  nl += 2;
  hip_check(hipGetLastError(), "inverse panel kernels");
  timing_end(tok, st, nl, g_flops_);
  nm.g_ready_ = true;
}

// The sweep kernels instantiated for the register tiles of a level (a power of two).
template <class T, class... A>
static void launch_fwd_task(int cols, dim3 grid, hipStream_t st, A... a) {
  switch (cols) {
    case 2: hipLaunchKernelGGL((k_fwd_task<T, 2>), grid, dim3(256), 0, st, a...); break;
    case 4: hipLaunchKernelGGL((k_fwd_task<T, 4>), grid, dim3(256), 0, st, a...); break;
    case 8: hipLaunchKernelGGL((k_fwd_task<T, 8>), grid, dim3(256), 0, st, a...); break;
    case 16: hipLaunchKernelGGL((k_fwd_task<T, 16>), grid, dim3(256), 0, st, a...); break;
    default: hipLaunchKernelGGL((k_fwd_task<T, kFwdCols>), grid, dim3(256), 0, st, a...); break;
  }
}
template <class T, class... A>
static void launch_fwd_wave(int cols, dim3 grid, hipStream_t st, A... a) {
  switch (cols) {
    case 8: hipLaunchKernelGGL((k_fwd_wave<T, 8>), grid, dim3(256), 0, st, a...); break;
    case 16: hipLaunchKernelGGL((k_fwd_wave<T, 16>), grid, dim3(256), 0, st, a...); break;
    default: hipLaunchKernelGGL((k_fwd_wave<T, kWaveCols>), grid, dim3(256), 0, st, a...); break;
  }
}
template <class T, class... A>
static void launch_bwd_task(int rows, dim3 grid, hipStream_t st, A... a) {
  switch (rows) {
    case 4: hipLaunchKernelGGL((k_bwd_task<T, 4>), grid, dim3(256), 0, st, a...); break;
    case 8: hipLaunchKernelGGL((k_bwd_task<T, 8>), grid, dim3(256), 0, st, a...); break;
    case 16: hipLaunchKernelGGL((k_bwd_task<T, 16>), grid, dim3(256), 0, st, a...); break;
    default: hipLaunchKernelGGL((k_bwd_task<T, 32>), grid, dim3(256), 0, st, a...); break;
  }
}

template <class T>
void Engine::solve_t(Numeric& nm, double* x, int64_t nrhs, int64_t ldx, hipStream_t st) {
  if (!nm.g_ready_) build_inverse_panels<T>(nm, st);
  DevSym S{sym_.n, sym_.nsn, d_first_, d_ncol_, d_m_, d_rowoff_, d_panel_off_, d_cb_off_, d_w_off_,
           d_child_ptr_, d_child_list_, d_relmap_, d_rows_orig_};
  DevNum<T> N{(T*)nm.d_L_, (T*)nm.d_CB_, (T*)nm.d_W_, (T*)nm.d_dval_, (T*)nm.d_doff_, nm.d_ptype_, nm.d_piv_, nm.d_ipiv_, nm.d_cnt_};
  if (nm.ws_cols_ < nrhs) {
    for (void* q : {nm.d_Xp_, nm.d_Y_, nm.d_U_, nm.d_part_}) dfree_after(q, nm.streams_);
    nm.d_Xp_ = nm.d_Y_ = nm.d_U_ = nm.d_part_ = nullptr;
    const int64_t nrc = (nrhs + kMaxRhsChunk - 1) / kMaxRhsChunk;
    dalloc_into(nm.d_Xp_, (size_t)sym_.n * nrhs * sizeof(T), "hipMalloc (solve workspace)");
    dalloc_into(nm.d_Y_, (size_t)sym_.n * nrhs * sizeof(T), "hipMalloc (solve workspace)");
    dalloc_into(nm.d_U_, (size_t)std::max<int64_t>(1, rows_total_) * nrhs * sizeof(T),
                "hipMalloc (solve workspace)");
    // two halves: a backward level writes its partials into one while the next level's launch
    // reads the previous level's from the other
    dalloc_into(nm.d_part_, (size_t)std::max<int64_t>(1, 2 * part_slots_ * nrc) * kPartSlot * sizeof(T),
                "hipMalloc (solve workspace)");
    nm.ws_cols_ = nrhs;
  }
  T* part = (T*)nm.d_part_;
  T* Xp = (T*)nm.d_Xp_;
  T* Y = (T*)nm.d_Y_;
  T* U = (T*)nm.d_U_;
  const T* G = (const T*)nm.d_G_;
  const T* Gt = (const T*)nm.d_Gt_;
  const int32_t* pos = nm.d_pos_;
  const int32_t* rpos = nm.d_rpos_;
  const int p = (int)nrhs;
  // algorithmic bytes per sweep: every stored factor entry once, the right-hand sides read and
  // written once (FP64)
  const double sweep_bytes = sym_.factor_entries * (double)sizeof(T) + 2.0 * sym_.n * (double)p * sizeof(double);
  hipLaunchKernelGGL((k_permute_in<double, T>), dim3(blocks_for(sym_.n)), dim3(256), 0, st, sym_.n, p,
                     (const double*)x, ldx, pos, Xp);
  int64_t tok = timing_begin(kPhaseForward, st);
  int64_t nl = 0;
  // forward: [y1; u] = G z1 + [0; u_children] per front, z gathered from the children where read
  const int32_t* cidx = (const int32_t*)d_cidx_;
  for (const Level& L : levels_) {
    if (L.ft_wave)
      launch_fwd_wave<T>(L.ft_cols, dim3((unsigned)((L.ft_count + 3) / 4)), st, d_ftasks_ + L.ft_begin, (int)L.ft_count,
                         p, G, (const T*)Xp, Y, U, (const PullRow*)nm.d_zpull_, (const PullRow*)d_cpull_, cidx);
    else
      launch_fwd_task<T>(L.ft_cols, dim3((unsigned)L.ft_count), st, d_ftasks_ + L.ft_begin, p, G, (const T*)Xp, Y, U,
                         part, (const PullRow*)nm.d_zpull_, (const PullRow*)d_cpull_, cidx);
    ++nl;
    if (L.fh_count > 0) {
      hipLaunchKernelGGL(k_fwd_reduce<T>, dim3((unsigned)L.fh_count), dim3(256), 0, st, d_fheads_ + L.fh_begin, p, Y,
                         U, (const T*)part, (const PullRow*)d_cpull_, cidx);
      ++nl;
    }
  }
  timing_end(tok, st, nl, sweep_bytes);
  hipLaunchKernelGGL(k_solve_diag<T>, dim3(blocks_for((int64_t)sym_.n * p)), dim3(256), 0, st, sym_.n, N.dval,
                     N.doff, N.ptype, Y, p);
  tok = timing_begin(kPhaseBackward, st);
  nl = 0;
  // backward: x1 = G^T [y1; x_cb] per front, top-down, into Xp
  // one launch per level: its tasks, plus one workgroup per split tile of the level above, which
  // adds that level's partials for the levels below (the level's own tasks add the ones they read)
  const int64_t half = part_slots_ * ((p + kMaxRhsChunk - 1) / kMaxRhsChunk) * (int64_t)kPartSlot;
  auto part_of = [&](int l) { return part + (l & 1) * half; };
  for (int l = (int)levels_.size() - 1; l >= 0; --l) {
    const Level& L = levels_[l];
    const Level* up = l + 1 < (int)levels_.size() ? &levels_[l + 1] : nullptr;
    const int64_t nh = up ? up->bh_count : 0;
    if (L.bt_count + nh > 0) {
      launch_bwd_task<T>(L.bt_regs, dim3((unsigned)(L.bt_count + nh)), st, Gt, d_btasks_ + L.bt_begin, (int)L.bt_count,
                         d_bheads_ + (up ? up->bh_begin : 0), p, (const T*)Y, Xp, rpos, (const PivSplit*)d_psplit_, l,
                         (const T*)part_of(l + 1), part_of(l));
      ++nl;
    }
  }
  if (!levels_.empty() && levels_[0].bh_count > 0) {  // split tiles of the lowest level
    hipLaunchKernelGGL(k_bwd_reduce<T>, dim3((unsigned)levels_[0].bh_count), dim3(256), 0, st,
                       d_bheads_ + levels_[0].bh_begin, p, Xp, (const T*)part_of(0));
    ++nl;
  }
  timing_end(tok, st, nl, sweep_bytes);
  hipLaunchKernelGGL((k_permute_out<T, double>), dim3(blocks_for(sym_.n)), dim3(256), 0, st, sym_.n, p, (const T*)Xp,
                     pos, x, ldx);
  hip_check(hipGetLastError(), "solve kernels");
}

void Engine::solve(Numeric& nm, double* x, int64_t nrhs, int64_t ldx, hipStream_t stream) {
  if (!nm.valid_) throw std::runtime_error("solve without a valid factorization");
  if (nrhs <= 0) return;
  if (ldx < sym_.n) throw std::runtime_error("solve: ldx < n");
  nm.note_stream(stream);
  if (nm.precision_ == 0) solve_t<double>(nm, x, nrhs, ldx, stream);
  else solve_t<float>(nm, x, nrhs, ldx, stream);
}


}  // namespace hsg
