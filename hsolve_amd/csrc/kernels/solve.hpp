// Multifrontal forward / backward substitution with p right-hand sides, level-batched and blocked
// (the reference's MUMPS JOB = 3, src/factor/mumps_factorization.cpp:458-527).
//
// Each front's pivot block L11 (unit lower, ncol x ncol) is split into column blocks of kSolveBlock;
// the inverses of their diagonal blocks are formed once after the factorization
// (k_block_inverse), so every step of a triangular solve is a small dense product:
//   forward, per front:  gather own rows (pivot order) and children's update vectors;
//                        for each block b: y_b = inv(L_bb) z_b (one workgroup per front),
//                        then every row below the block - remaining pivot rows and contribution
//                        rows alike - is updated, z_i -= L(i, b) y_b (many workgroups per front);
//                        the contribution rows end as the update vector U passed to the parent;
//   diagonal:            z = D^-1 z (1x1 and 2x2 blocks);
//   backward, per front: z_j -= sum_i L21(i, j) x_cb(i) with the ancestors' rows (final already),
//                        then for each block b from the last: x_b = inv(L_bb)^T z_b, and
//                        z_j -= sum_{i in b} L(i, j) x_i for the pivot columns j before the block;
//                        scatter into X.
// Every launch covers all fronts of one level (grid y = front). R is the right-hand-side type
// (double), T the factor type.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "front_common.hpp"

namespace hsg {

constexpr int kSolveBlock = 64;
constexpr int kSolveRows = 64;   // rows per update workgroup
constexpr int kMaxRhsChunk = 8;  // right-hand sides per pass

// L(i, j) of a pivot block, 0 at the (j+1, j) position of a 2x2 or null pivot pair.
template <class T>
__device__ __forceinline__ T lpiv_at(const T* P, int m, int i, int j, const int8_t* pt) {
  if (i == j + 1 && (pt[j] == kPiv2First || pt[j] == kPivNull)) return (T)0;
  return P[(int64_t)j * m + i];
}

// Inverse of the unit lower diagonal block b of front s (bs x bs, column-major, into Linv).
// One workgroup per (front, block) pair of `pairs` (2 ints each).
template <class T>
__global__ void __launch_bounds__(256) k_block_inverse(DevSym S, DevNum<T> N, const int32_t* __restrict__ pairs,
                                                        const int64_t* __restrict__ linv_off, T* __restrict__ Linv) {
  __shared__ T Lb[kSolveBlock][kSolveBlock + 1];
  const int s = pairs[2 * blockIdx.x], b = pairs[2 * blockIdx.x + 1];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int r0 = b * kSolveBlock;
  const int bs = min(kSolveBlock, ncol - r0);
  const T* P = N.L + S.panel_off[s];
  const int8_t* pt = N.ptype + first;
  for (int idx = threadIdx.x; idx < bs * bs; idx += blockDim.x) {
    const int i = idx % bs, j = idx / bs;
    Lb[i][j] = i > j ? lpiv_at(P, m, r0 + i, r0 + j, pt) : (T)0;
  }
  __syncthreads();
  __shared__ T Xi[kSolveBlock][kSolveBlock + 1];
  T* out = Linv + linv_off[s] + (int64_t)b * kSolveBlock * kSolveBlock;
  const int j = threadIdx.x;
  if (j < bs) {
    // column j of the inverse: x_i = delta_ij - sum_{j <= t < i} L(i, t) x_t
    for (int i = 0; i < j; ++i) Xi[i][j] = (T)0;
    for (int i = j; i < bs; ++i) {
      T v = i == j ? (T)1 : (T)0;
      for (int t = j; t < i; ++t) v -= Lb[i][t] * Xi[t][j];
      Xi[i][j] = v;
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < bs * bs; idx += blockDim.x) out[idx] = Xi[idx % bs][idx / bs];
}

// ---------------------------------------------------------------------------------- forward

template <class T, class R>
__global__ void __launch_bounds__(256) k_fwd_gather(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                     const R* __restrict__ X, int64_t ldx, int p, T* Z, T* U,
                                                     int64_t ldu) {
  const int s = list[blockIdx.x];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int cb = m - ncol;
  const int n = S.n;
  const int64_t roff = S.sn_rowoff[s];
  const int32_t* rows = S.rows_orig + roff;
  T* zs = Z + first;
  T* us = U + roff + ncol;
  for (int idx = threadIdx.x; idx < ncol * p; idx += blockDim.x) {
    const int i = idx % ncol, c = idx / ncol;
    zs[i + (int64_t)c * n] = (T)X[rows[N.piv[first + i]] + (int64_t)c * ldx];
  }
  for (int idx = threadIdx.x; idx < cb * p; idx += blockDim.x) {
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
    for (int idx = threadIdx.x; idx < cbc * p; idx += blockDim.x) {
      const int i = idx % cbc, c = idx / cbc;
      const int loc = rel[i];
      const T v = uc[i + (int64_t)c * ldu];
      if (loc < ncol) zs[N.ipiv[first + loc] + (int64_t)c * n] += v;
      else us[loc - ncol + (int64_t)c * ldu] += v;
    }
    __syncthreads();
  }
}

// y_b = inv(L_bb) z_b in place (forward) or x_b = inv(L_bb)^T z_b (backward); one workgroup per front.
template <class T, bool kTranspose>
__global__ void __launch_bounds__(256) k_diag_block(DevSym S, DevNum<T> N, const int32_t* __restrict__ list, int b,
                                                     const int64_t* __restrict__ linv_off, const T* __restrict__ Linv,
                                                     int p, T* Z) {
  __shared__ T Li[kSolveBlock][kSolveBlock + 1];
  __shared__ T zb[kSolveBlock][kMaxRhsChunk];
  const int s = list[blockIdx.x];
  const int ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int r0 = b * kSolveBlock;
  if (r0 >= ncol) return;
  const int bs = min(kSolveBlock, ncol - r0);
  const int n = S.n;
  const T* li = Linv + linv_off[s] + (int64_t)b * kSolveBlock * kSolveBlock;
  for (int idx = threadIdx.x; idx < bs * bs; idx += blockDim.x) Li[idx % bs][idx / bs] = li[idx];
  T* zs = Z + first + r0;
  for (int c0 = 0; c0 < p; c0 += kMaxRhsChunk) {
    const int pc = min(kMaxRhsChunk, p - c0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < bs * pc; idx += blockDim.x) {
      const int i = idx % bs, c = idx / bs;
      zb[i][c] = zs[i + (int64_t)(c0 + c) * n];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < bs * pc; idx += blockDim.x) {
      const int i = idx % bs, c = idx / bs;
      T acc = (T)0;
      if (!kTranspose) {
        for (int t = 0; t <= i; ++t) acc += Li[i][t] * zb[t][c];
      } else {
        for (int t = i; t < bs; ++t) acc += Li[t][i] * zb[t][c];
      }
      zs[i + (int64_t)(c0 + c) * n] = acc;
    }
  }
}

// Rows below block b: z_i -= L(i, b) y_b for pivot rows, u_i -= ... for contribution rows.
// grid.x: chunks of kSolveRows rows over [r0 + bs, m); 256 threads = kSolveRows rows x 4 groups of
// the block's columns, partial sums combined in LDS.
template <class T>
__global__ void __launch_bounds__(256) k_fwd_update(DevSym S, DevNum<T> N, const int32_t* __restrict__ list, int b,
                                                     int p, T* Z, T* U, int64_t ldu) {
  __shared__ T yb[kSolveBlock][kMaxRhsChunk];
  __shared__ T part[4][kSolveRows][kMaxRhsChunk + 1];
  const int s = list[blockIdx.y];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int r0 = b * kSolveBlock;
  if (r0 >= ncol) return;
  const int bs = min(kSolveBlock, ncol - r0);
  const int rstart = r0 + bs + blockIdx.x * kSolveRows;
  if (rstart >= m) return;
  const int rn = min(kSolveRows, m - rstart);
  const int n = S.n;
  const T* P = N.L + S.panel_off[s];
  const int8_t* pt = N.ptype + first;
  T* zs = Z + first;
  T* us = U + S.sn_rowoff[s] + ncol;
  const int lr = threadIdx.x % kSolveRows, g = threadIdx.x / kSolveRows;
  for (int c0 = 0; c0 < p; c0 += kMaxRhsChunk) {
    const int pc = min(kMaxRhsChunk, p - c0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < bs * pc; idx += blockDim.x) {
      const int t = idx % bs, c = idx / bs;
      yb[t][c] = zs[r0 + t + (int64_t)(c0 + c) * n];
    }
    __syncthreads();
    T acc[kMaxRhsChunk];
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = (T)0;
    const int i = rstart + lr;
    if (lr < rn) {
      for (int t = g; t < bs; t += 4) {
        const T l = lpiv_at(P, m, i, r0 + t, pt);
#pragma unroll
        for (int c = 0; c < kMaxRhsChunk; ++c)
          if (c < pc) acc[c] += l * yb[t][c];
      }
    }
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) part[g][lr][c] = acc[c];
    __syncthreads();
    if (g == 0 && lr < rn) {
      for (int c = 0; c < pc; ++c) {
        const T v = part[0][lr][c] + part[1][lr][c] + part[2][lr][c] + part[3][lr][c];
        if (i < ncol) zs[i + (int64_t)(c0 + c) * n] -= v;
        else us[i - ncol + (int64_t)(c0 + c) * ldu] -= v;
      }
    }
  }
}

// Fused forward block step b: every workgroup forms y_b = inv(L_bb) z_b for itself (small, from
// LDS), workgroup 0 stores it into Y (the forward result, a separate buffer, so that no
// workgroup of this launch reads a value another one rewrites), and each workgroup updates its
// chunk of the rows below the block: z_i -= L(i, b) y_b (pivot rows in Z, contribution rows in U).
template <class T>
__global__ void __launch_bounds__(256) k_fwd_step(DevSym S, DevNum<T> N, const int32_t* __restrict__ list, int b,
                                                   const int64_t* __restrict__ linv_off, const T* __restrict__ Linv,
                                                   int p, T* Z, T* Y, T* U, int64_t ldu) {
  __shared__ T Li[kSolveBlock][kSolveBlock + 1];
  __shared__ T zb[kSolveBlock][kMaxRhsChunk];
  __shared__ T yb[kSolveBlock][kMaxRhsChunk];
  __shared__ T part[4][kSolveRows][kMaxRhsChunk + 1];
  const int s = list[blockIdx.y];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int r0 = b * kSolveBlock;
  if (r0 >= ncol) return;
  const int bs = min(kSolveBlock, ncol - r0);
  const int rstart = r0 + bs + blockIdx.x * kSolveRows;
  if (blockIdx.x > 0 && rstart >= m) return;
  const int rn = max(0, min(kSolveRows, m - rstart));
  const int n = S.n;
  const T* P = N.L + S.panel_off[s];
  const int8_t* pt = N.ptype + first;
  T* zs = Z + first;
  T* ys = Y + first;
  T* us = U + S.sn_rowoff[s] + ncol;
  const T* li = Linv + linv_off[s] + (int64_t)b * kSolveBlock * kSolveBlock;
  for (int idx = threadIdx.x; idx < bs * bs; idx += blockDim.x) Li[idx % bs][idx / bs] = li[idx];
  const int lr = threadIdx.x % kSolveRows, g = threadIdx.x / kSolveRows;
  for (int c0 = 0; c0 < p; c0 += kMaxRhsChunk) {
    const int pc = min(kMaxRhsChunk, p - c0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < bs * pc; idx += blockDim.x) {
      const int t = idx % bs, c = idx / bs;
      zb[t][c] = zs[r0 + t + (int64_t)(c0 + c) * n];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < bs * pc; idx += blockDim.x) {
      const int i = idx % bs, c = idx / bs;
      T acc = (T)0;
      for (int t = 0; t <= i; ++t) acc += Li[i][t] * zb[t][c];
      yb[i][c] = acc;
      if (blockIdx.x == 0) ys[r0 + i + (int64_t)(c0 + c) * n] = acc;
    }
    __syncthreads();
    T acc[kMaxRhsChunk];
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = (T)0;
    const int i = rstart + lr;
    if (lr < rn) {
      for (int t = g; t < bs; t += 4) {
        const T l = lpiv_at(P, m, i, r0 + t, pt);
#pragma unroll
        for (int c = 0; c < kMaxRhsChunk; ++c)
          if (c < pc) acc[c] += l * yb[t][c];
      }
    }
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) part[g][lr][c] = acc[c];
    __syncthreads();
    if (g == 0 && lr < rn) {
      for (int c = 0; c < pc; ++c) {
        const T v = part[0][lr][c] + part[1][lr][c] + part[2][lr][c] + part[3][lr][c];
        if (i < ncol) zs[i + (int64_t)(c0 + c) * n] -= v;
        else us[i - ncol + (int64_t)(c0 + c) * ldu] -= v;
      }
    }
  }
}

// Fused backward block step b: every workgroup forms x_b = inv(L_bb)^T y_b, workgroup 0 stores it
// into X (the backward result in pivot order, a separate buffer), and the workgroups of the
// column chunks before the block update y_j -= sum_{i in b} L(i, j) x_i for their columns.
// grid.x = 1 + number of 64-column chunks in [0, r0).
template <class T>
__global__ void __launch_bounds__(256) k_bwd_step(DevSym S, DevNum<T> N, const int32_t* __restrict__ list, int b,
                                                   const int64_t* __restrict__ linv_off, const T* __restrict__ Linv,
                                                   int p, T* Y, T* Xo) {
  __shared__ T Li[kSolveBlock][kSolveBlock + 1];
  __shared__ T Lt[kSolveBlock][kSolveBlock + 1];
  __shared__ T yb[kSolveBlock][kMaxRhsChunk];
  __shared__ T xb[kSolveBlock][kMaxRhsChunk];
  const int s = list[blockIdx.y];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int r0 = b * kSolveBlock;
  if (r0 >= ncol) return;
  const int bs = min(kSolveBlock, ncol - r0);
  const int chunk = (int)blockIdx.x - 1;  // -1: the storing workgroup
  const int j0 = chunk * kSolveBlock;
  if (chunk >= 0 && j0 >= r0) return;
  const int jn = chunk >= 0 ? min(kSolveBlock, r0 - j0) : 0;
  const int n = S.n;
  const T* P = N.L + S.panel_off[s];
  const int8_t* pt = N.ptype + first;
  T* ys = Y + first;
  T* xs = Xo + first;
  const T* li = Linv + linv_off[s] + (int64_t)b * kSolveBlock * kSolveBlock;
  for (int idx = threadIdx.x; idx < bs * bs; idx += blockDim.x) Li[idx % bs][idx / bs] = li[idx];
  if (chunk >= 0)
    for (int idx = threadIdx.x; idx < kSolveBlock * kSolveBlock; idx += blockDim.x) {
      const int t = idx % kSolveBlock, j = idx / kSolveBlock;
      Lt[t][j] = (t < bs && j < jn) ? lpiv_at(P, m, r0 + t, j0 + j, pt) : (T)0;
    }
  const int jj = threadIdx.x % kSolveBlock, cq = threadIdx.x / kSolveBlock;
  for (int c0 = 0; c0 < p; c0 += kMaxRhsChunk) {
    const int pc = min(kMaxRhsChunk, p - c0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < bs * pc; idx += blockDim.x) {
      const int t = idx % bs, c = idx / bs;
      yb[t][c] = ys[r0 + t + (int64_t)(c0 + c) * n];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < kSolveBlock * kMaxRhsChunk; idx += blockDim.x) {
      const int i = idx % kSolveBlock, c = idx / kSolveBlock;
      T acc = (T)0;
      if (i < bs && c < pc)
        for (int t = i; t < bs; ++t) acc += Li[t][i] * yb[t][c];
      xb[i][c] = acc;
      if (chunk < 0 && i < bs && c < pc) xs[r0 + i + (int64_t)(c0 + c) * n] = acc;
    }
    __syncthreads();
    if (chunk >= 0) {
      T acc0 = (T)0, acc1 = (T)0;
      for (int t = 0; t < bs; ++t) {
        const T l = Lt[t][jj];
        acc0 += l * xb[t][cq];
        acc1 += l * xb[t][cq + 4];
      }
      if (jj < jn) {
        if (cq < pc) ys[j0 + jj + (int64_t)(c0 + cq) * n] -= acc0;
        if (cq + 4 < pc) ys[j0 + jj + (int64_t)(c0 + cq + 4) * n] -= acc1;
      }
    }
  }
}

// ---------------------------------------------------------------------------------- diagonal

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
      const T a = dval[i], bb = doff[i], d = dval[i + 1];
      const T det = a * d - bb * bb;
      const T z1 = z[i], z2 = z[i + 1];
      z[i] = (d * z1 - bb * z2) / det;
      z[i + 1] = (a * z2 - bb * z1) / det;
    } else if (t == kPivNull) {
      z[i] = (T)0;
    }
  }
}

// ---------------------------------------------------------------------------------- backward

// x_cb = X at the front's contribution rows (ancestors' rows, final), into the U buffer.
template <class T, class R>
__global__ void __launch_bounds__(256) k_bwd_gather(DevSym S, const int32_t* __restrict__ list, const R* X,
                                                     int64_t ldx, int p, T* U, int64_t ldu) {
  const int s = list[blockIdx.x];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s];
  const int cb = m - ncol;
  const int64_t off = S.sn_rowoff[s] + ncol;
  const int32_t* rows = S.rows_orig + off;
  for (int idx = threadIdx.x; idx < cb * p; idx += blockDim.x) {
    const int i = idx % cb, c = idx / cb;
    U[off + i + (int64_t)c * ldu] = (T)X[rows[i] + (int64_t)c * ldx];
  }
}

// z_j -= sum_i L21(i, j) x_cb(i) for the pivot columns j of chunk blockIdx.x (64 columns):
// tiles of 64 contribution rows x 64 columns staged through LDS with coalesced column reads,
// one thread per (column, right-hand-side group) accumulating over the rows.
template <class T>
__global__ void __launch_bounds__(256) k_bwd_cb(DevSym S, DevNum<T> N, const int32_t* __restrict__ list, int p, T* Z,
                                                 const T* U, int64_t ldu) {
  __shared__ T Lt[kSolveBlock][kSolveBlock + 1];
  __shared__ T xt[kSolveBlock][kMaxRhsChunk];
  const int s = list[blockIdx.y];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int cb = m - ncol;
  const int j0 = blockIdx.x * kSolveBlock;
  if (cb == 0 || j0 >= ncol) return;
  const int n = S.n;
  const T* P = N.L + S.panel_off[s];
  const T* xc = U + S.sn_rowoff[s] + ncol;
  T* zs = Z + first;
  const int jj = threadIdx.x % kSolveBlock, cq = threadIdx.x / kSolveBlock;  // cq in [0, 4)
  const int jn = min(kSolveBlock, ncol - j0);
  for (int c0 = 0; c0 < p; c0 += kMaxRhsChunk) {
    const int pc = min(kMaxRhsChunk, p - c0);
    T acc0 = (T)0, acc1 = (T)0;
    for (int r0 = 0; r0 < cb; r0 += kSolveBlock) {
      const int rn = min(kSolveBlock, cb - r0);
      __syncthreads();
      for (int idx = threadIdx.x; idx < kSolveBlock * kSolveBlock; idx += blockDim.x) {
        const int r = idx % kSolveBlock, j = idx / kSolveBlock;
        Lt[r][j] = (r < rn && j < jn) ? P[(int64_t)(j0 + j) * m + ncol + r0 + r] : (T)0;
      }
      for (int idx = threadIdx.x; idx < kSolveBlock * kMaxRhsChunk; idx += blockDim.x) {
        const int r = idx % kSolveBlock, c = idx / kSolveBlock;
        xt[r][c] = (r < rn && c < pc) ? xc[r0 + r + (int64_t)(c0 + c) * ldu] : (T)0;
      }
      __syncthreads();
      for (int r = 0; r < rn; ++r) {
        const T l = Lt[r][jj];
        acc0 += l * xt[r][cq];
        acc1 += l * xt[r][cq + 4];
      }
    }
    if (jj < jn) {
      if (cq < pc) zs[j0 + jj + (int64_t)(c0 + cq) * n] -= acc0;
      if (cq + 4 < pc) zs[j0 + jj + (int64_t)(c0 + cq + 4) * n] -= acc1;
    }
  }
}

// z_j -= sum_{i in block b} L(i, j) x_i for pivot columns j < r0 (chunks of 64 columns); the
// block's rows of those columns are staged through LDS, one thread per (column, rhs group).
template <class T>
__global__ void __launch_bounds__(256) k_bwd_update(DevSym S, DevNum<T> N, const int32_t* __restrict__ list, int b,
                                                     int p, T* Z) {
  __shared__ T Lt[kSolveBlock][kSolveBlock + 1];
  __shared__ T xb[kSolveBlock][kMaxRhsChunk];
  const int s = list[blockIdx.y];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int r0 = b * kSolveBlock;
  if (r0 >= ncol) return;
  const int bs = min(kSolveBlock, ncol - r0);
  const int j0 = blockIdx.x * kSolveBlock;
  if (j0 >= r0) return;
  const int jn = min(kSolveBlock, r0 - j0);
  const int n = S.n;
  const T* P = N.L + S.panel_off[s];
  const int8_t* pt = N.ptype + first;
  T* zs = Z + first;
  const int jj = threadIdx.x % kSolveBlock, cq = threadIdx.x / kSolveBlock;
  for (int idx = threadIdx.x; idx < kSolveBlock * kSolveBlock; idx += blockDim.x) {
    const int t = idx % kSolveBlock, j = idx / kSolveBlock;
    Lt[t][j] = (t < bs && j < jn) ? lpiv_at(P, m, r0 + t, j0 + j, pt) : (T)0;
  }
  for (int c0 = 0; c0 < p; c0 += kMaxRhsChunk) {
    const int pc = min(kMaxRhsChunk, p - c0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < kSolveBlock * kMaxRhsChunk; idx += blockDim.x) {
      const int t = idx % kSolveBlock, c = idx / kSolveBlock;
      xb[t][c] = (t < bs && c < pc) ? zs[r0 + t + (int64_t)(c0 + c) * n] : (T)0;
    }
    __syncthreads();
    T acc0 = (T)0, acc1 = (T)0;
    for (int t = 0; t < bs; ++t) {
      const T l = Lt[t][jj];
      acc0 += l * xb[t][cq];
      acc1 += l * xb[t][cq + 4];
    }
    if (jj < jn) {
      if (cq < pc) zs[j0 + jj + (int64_t)(c0 + cq) * n] -= acc0;
      if (cq + 4 < pc) zs[j0 + jj + (int64_t)(c0 + cq + 4) * n] -= acc1;
    }
  }
}

template <class T, class R>
__global__ void __launch_bounds__(256) k_bwd_scatter(DevSym S, DevNum<T> N, const int32_t* __restrict__ list, R* X,
                                                      int64_t ldx, int p, const T* Z) {
  const int s = list[blockIdx.x];
  const int ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int n = S.n;
  const int32_t* rows = S.rows_orig + S.sn_rowoff[s];
  for (int idx = threadIdx.x; idx < ncol * p; idx += blockDim.x) {
    const int i = idx % ncol, c = idx / ncol;
    X[rows[N.piv[first + i]] + (int64_t)c * ldx] = (R)Z[first + i + (int64_t)c * n];
  }
}

}  // namespace hsg
