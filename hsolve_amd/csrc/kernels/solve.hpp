// Multifrontal forward / backward substitution with p right-hand sides in partitioned-inverse
// form (the solve phase the reference hands to MUMPS, JOB = 3,
// src/factor/mumps_factorization.cpp:458-527).
//
// With P K P^T = L D L^T and, per front s (ncol pivots, m rows), L = [L11; L21], the solve phase
// reads, per front,
//     G  = [ G11 ; G21 ] = [ L11^-1 ; -L21 L11^-1 ]   (m x ncol, column-major, the panel layout)
//     Gt = G stored row-major                         (same offsets, element (i, j) at i*ncol + j)
// because L^-1 restricted to one front is [L11^-1 0; -L21 L11^-1 I]. Then
//   forward, bottom-up per level:  z = own right-hand-side rows + children's update vectors
//                                  (pull gather, fixed order), then one GEMV per front:
//                                  [ y1 ; u ] = G z1 + [ 0 ; z_cb ]
//   diagonal:                      y = D^-1 y (1x1 and 2x2 blocks)
//   backward, top-down per level:  x1 = G^T [ y1 ; x_cb ] with x_cb read from X (rows of
//                                  ancestors, final), one GEMV per front, scattered into X.
// Two launches per level forward, one backward, and no sequential chain inside a front.
//
// G is built from L once per factorization by batched 64 x 64 tile products: inverses of the
// 64 x 64 diagonal blocks of L11, then recursive doubling over block sizes h = 64, 128, ...
//     inv([A 0; C B]) = [inv(A) 0; -inv(B) C inv(A)  inv(B)],
// then G21 = -L21 G11, then the row-major copy.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "front_common.hpp"

namespace hsg {

constexpr int kSolveBlock = 64;   // diagonal block of the first inversion step
constexpr int kGvTile = 64;       // rows (forward) / columns (backward) per GEMV workgroup
constexpr int kGvChunk = 128;     // vector entries staged in LDS per pass
constexpr int kMaxRhsChunk = 8;   // right-hand sides per pass
constexpr int kGatherRows = 256;  // rows per gather workgroup

// L(i, j) of a pivot block, 0 at the (j+1, j) position of a 2x2 or null pivot pair (that slot of
// the panel holds D's off-diagonal).
template <class T>
__device__ __forceinline__ T lpiv_at(const T* P, int m, int i, int j, const int8_t* pt) {
  if (i == j + 1 && (pt[j] == kPiv2First || pt[j] == kPivNull)) return (T)0;
  return P[(int64_t)j * m + i];
}

// ------------------------------------------------------------------ building G from L

// Inverse of the unit lower diagonal block b of front s into G (bs x bs, zeros above the
// diagonal), and the (j+1, j) slots of the block's pivot pairs in L set to 0 so that the tile
// products below read L as it is. One 64-thread workgroup per (front, block) pair of `pairs`.
// Outer-product form: X = I, then for t = 0.. each row i > t does X(i, 0:t) -= L(i, t) X(t, 0:t).
template <class T>
__global__ void __launch_bounds__(64) k_diag_block_inverse(DevSym S, DevNum<T> N, const int32_t* __restrict__ pairs,
                                                            T* __restrict__ G) {
  __shared__ T Xi[kSolveBlock * (kSolveBlock + 1)];
  __shared__ T rowt[kSolveBlock];
  constexpr int ld = kSolveBlock + 1;
  const int s = pairs[2 * blockIdx.x], b = pairs[2 * blockIdx.x + 1];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int r0 = b * kSolveBlock;
  const int bs = min(kSolveBlock, ncol - r0);
  T* P = N.L + S.panel_off[s];
  const int8_t* pt = N.ptype + first;
  const int lane = threadIdx.x;
  for (int j = 0; j < bs; ++j) Xi[lane + j * ld] = lane == j ? (T)1 : (T)0;
  __syncthreads();
  for (int t = 0; t + 1 < bs; ++t) {
    rowt[lane] = lane <= t ? Xi[t + lane * ld] : (T)0;
    __syncthreads();
    const T lit = (lane > t && lane < bs) ? lpiv_at(P, m, r0 + lane, r0 + t, pt) : (T)0;
    if (lit != (T)0) {
      T* row = Xi + lane;
#pragma unroll 8
      for (int j = 0; j <= t; ++j) row[j * ld] -= lit * rowt[j];
    }
    __syncthreads();
  }
  T* Gs = G + S.panel_off[s];
  if (lane < bs)
    for (int j = 0; j < bs; ++j) Gs[(int64_t)(r0 + j) * m + r0 + lane] = Xi[lane + j * ld];
  // pivot pairs whose second row lies inside this front
  if (lane < bs) {
    const int j = r0 + lane;
    if (j + 1 < ncol && (pt[j] == kPiv2First || pt[j] == kPivNull)) P[(int64_t)j * m + j + 1] = (T)0;
  }
}

// C = (neg ? -1 : 1) * A(:, k0:k1) B(k0:k1, :) for one tile of at most 64 x 64, column-major
// operands given as element offsets into one of three buffers (0: L, 1: G, 2: scratch).
struct GemmTask {
  int64_t a, b, c;
  int32_t lda, ldb, ldc;
  int32_t rows, cols, k0, k1;
  int8_t abuf, bbuf, cbuf, neg;
};

constexpr int kGmTile = 64;
constexpr int kGmK = 16;

template <class T>
__device__ __forceinline__ T* pick_buffer(int which, T* b0, T* b1, T* b2) {
  return which == 0 ? b0 : (which == 1 ? b1 : b2);
}

template <class T>
__global__ void __launch_bounds__(256) k_tile_gemm(const GemmTask* __restrict__ tasks, T* b0, T* b1, T* b2) {
  const GemmTask tk = tasks[blockIdx.x];
  const T* A = pick_buffer<T>(tk.abuf, b0, b1, b2) + tk.a;
  const T* B = pick_buffer<T>(tk.bbuf, b0, b1, b2) + tk.b;
  T* C = pick_buffer<T>(tk.cbuf, b0, b1, b2) + tk.c;
  __shared__ T As[kGmK][kGmTile + 1];  // As[k][i]
  __shared__ T Bs[kGmK][kGmTile + 1];  // Bs[k][j]
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (T)0;
  for (int kk = tk.k0; kk < tk.k1; kk += kGmK) {
    for (int idx = tid; idx < kGmK * kGmTile; idx += 256) {
      const int r = idx % kGmTile, t = idx / kGmTile;  // A: consecutive threads along rows
      const int k = kk + t;
      As[t][r] = (r < tk.rows && k < tk.k1) ? A[(int64_t)k * tk.lda + r] : (T)0;
      const int t2 = idx % kGmK, c2 = idx / kGmK;      // B: consecutive threads along k
      const int k2 = kk + t2;
      Bs[t2][c2] = (c2 < tk.cols && k2 < tk.k1) ? B[(int64_t)c2 * tk.ldb + k2] : (T)0;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kGmK; ++t) {
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
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = tx + 16 * a, j = ty + 16 * b;
      if (i < tk.rows && j < tk.cols) C[(int64_t)j * tk.ldc + i] = tk.neg ? -acc[a][b] : acc[a][b];
    }
}

// Row-major copy of G: Gt(i, j) = G(i, j) on and below the diagonal of G11 and on all of G21,
// 0 above the diagonal. One workgroup per 64 x 64 tile (front, tile row, tile column) of `tiles`.
template <class T>
__global__ void __launch_bounds__(256) k_transpose_g(DevSym S, const int32_t* __restrict__ tiles, const T* __restrict__ G,
                                                      T* __restrict__ Gt) {
  __shared__ T tl[64][65];
  const int s = tiles[3 * blockIdx.x], ti = tiles[3 * blockIdx.x + 1], tj = tiles[3 * blockIdx.x + 2];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s];
  const int64_t off = S.panel_off[s];
  const int i0 = ti * 64, j0 = tj * 64;
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int i = idx % 64, j = idx / 64;  // read along i (column-major G)
    const int gi = i0 + i, gj = j0 + j;
    T v = (T)0;
    if (gi < m && gj < ncol && gi >= gj) v = G[off + (int64_t)gj * m + gi];
    tl[i][j] = v;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int j = idx % 64, i = idx / 64;  // write along j (row-major Gt)
    const int gi = i0 + i, gj = j0 + j;
    if (gi < m && gj < ncol) Gt[off + (int64_t)gi * ncol + gj] = tl[i][j];
  }
}

// ---------------------------------------------------------------------------------- forward

// z of each front of the level: own rows take the right-hand side (converted to the factor type),
// every row adds the update vectors of the children's contribution rows mapped onto it (cptr /
// cidx: per front row, the U rows of the children, in child order). Pivot rows go to Z in pivot
// order, contribution rows to the front's U rows. grid (row chunks, fronts).
template <class T, class R>
__global__ void __launch_bounds__(kGatherRows) k_fwd_gather(DevSym S, DevNum<T> N, const int32_t* __restrict__ list,
                                                             const R* __restrict__ X, int64_t ldx, int p, T* Z, T* U,
                                                             int64_t ldu, const int64_t* __restrict__ cptr,
                                                             const int32_t* __restrict__ cidx) {
  const int s = list[blockIdx.y];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int loc = blockIdx.x * kGatherRows + threadIdx.x;
  if (loc >= m) return;
  const int64_t roff = S.sn_rowoff[s];
  const int64_t q0 = cptr[roff + loc], q1 = cptr[roff + loc + 1];
  const bool own = loc < ncol;
  const int64_t xrow = own ? S.rows_orig[roff + loc] : 0;
  const int64_t dst = own ? (int64_t)first + N.ipiv[first + loc] : roff + loc;
  T* out = own ? Z : U;
  const int64_t ldo = own ? (int64_t)S.n : ldu;
  for (int c = 0; c < p; ++c) {
    T v = own ? (T)X[xrow + c * ldx] : (T)0;
    for (int64_t q = q0; q < q1; ++q) v += U[cidx[q] + c * ldu];
    out[dst + c * ldo] = v;
  }
}

// Split products: a tile's long dimension (columns of G forward, rows of Gt backward) is cut into
// chunks of kGvSplit, one workgroup each. With more than one chunk, each workgroup stores its
// partial sums (64 x p) in its slot of `part`, and the last workgroup of the tile to finish
// (device-scope fence, then a counter) adds the partials in chunk order and writes the result,
// so the sum does not depend on the order the workgroups run in. The counter is reset by that
// workgroup for the next use. Counters: ((front index in level) * tiles + tile) * rhs_chunks +
// rhs_chunk (rhs_chunk = which group of kMaxRhsChunk right-hand sides); slots: counter * chunks +
// chunk.
constexpr int kGvSplit = 256;

// Returns true in the workgroup that holds the complete sums of the tile in red[.][lane][c]
// (red = this workgroup's own sums when the tile has one chunk).
template <class T>
__device__ __forceinline__ bool split_combine(T (*red)[kGvTile][kMaxRhsChunk + 1], int nchunk, int chunk,
                                              int64_t slot0, int pc, T* __restrict__ part, int* __restrict__ counter,
                                              int cidx) {
  const int lane = threadIdx.x % kGvTile, g = threadIdx.x / kGvTile;
  // red[0..3] hold the 4 group sums; fold them into red[0]
  for (int c = g; c < pc; c += 4) red[0][lane][c] = red[0][lane][c] + red[1][lane][c] + red[2][lane][c] + red[3][lane][c];
  __syncthreads();
  if (nchunk == 1) return true;
  T* mine = part + (slot0 + chunk) * (int64_t)(kGvTile * kMaxRhsChunk);
  for (int c = g; c < pc; c += 4) mine[c * kGvTile + lane] = red[0][lane][c];
  __threadfence();
  __syncthreads();
  __shared__ int s_last;
  if (threadIdx.x == 0) s_last = atomicAdd(&counter[cidx], 1) == nchunk - 1;
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  for (int c = g; c < pc; c += 4) {
    T v = (T)0;
    for (int q = 0; q < nchunk; ++q) v += part[(slot0 + q) * (int64_t)(kGvTile * kMaxRhsChunk) + c * kGvTile + lane];
    red[0][lane][c] = v;
  }
  if (threadIdx.x == 0) counter[cidx] = 0;
  __syncthreads();
  return true;
}

// [y1; u] = G z1 + [0; u] for a tile of 64 front rows and one chunk of G's columns: rows < ncol
// (G11, lower triangular, so only columns up to the tile's end) give y1 in Y, contribution rows
// add G21 z1 to the front's U rows. 256 threads = 64 rows x 4 column groups; z1 is staged through
// LDS, G read along columns. grid (row tiles, fronts, column chunks).
template <class T>
__global__ void __launch_bounds__(256) k_fwd_gemv(DevSym S, const T* __restrict__ G, const int32_t* __restrict__ list,
                                                   int p, const T* __restrict__ Z, T* __restrict__ Y, T* __restrict__ U,
                                                   int64_t ldu, T* __restrict__ part, int* __restrict__ counter) {
  __shared__ T zs[kGvChunk][kMaxRhsChunk];
  __shared__ T red[4][kGvTile][kMaxRhsChunk + 1];
  const int f = blockIdx.y;
  const int s = list[f];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int tile = blockIdx.x, chunk = blockIdx.z;
  const int r0 = tile * kGvTile;
  if (r0 >= m) return;
  const int kend = min(ncol, r0 + kGvTile);
  const int nchunk = (kend + kGvSplit - 1) / kGvSplit;
  if (chunk >= nchunk) return;
  const int kbeg = chunk * kGvSplit, kstop = min(kend, kbeg + kGvSplit);
  const int n = S.n;
  const T* Gs = G + S.panel_off[s];
  const int lane = threadIdx.x % kGvTile, g = threadIdx.x / kGvTile;
  const int i = r0 + lane;
  const bool rin = i < m;
  const T* zf = Z + first;
  const int nrc = (p + kMaxRhsChunk - 1) / kMaxRhsChunk;
  for (int c0 = 0; c0 < p; c0 += kMaxRhsChunk) {
    const int pc = min(kMaxRhsChunk, p - c0);
    const int cidx = (f * (int)gridDim.x + tile) * nrc + c0 / kMaxRhsChunk;
    const int64_t slot0 = (int64_t)cidx * gridDim.z;
    T acc[kMaxRhsChunk];
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = (T)0;
    for (int k0 = kbeg; k0 < kstop; k0 += kGvChunk) {
      const int kn = min(kGvChunk, kstop - k0);
      __syncthreads();
      for (int idx = threadIdx.x; idx < kGvChunk * kMaxRhsChunk; idx += 256) {
        const int t = idx % kGvChunk, c = idx / kGvChunk;
        zs[t][c] = (t < kn && c < pc) ? zf[k0 + t + (int64_t)(c0 + c) * n] : (T)0;
      }
      __syncthreads();
      if (rin) {
        const T* gcol = Gs + (int64_t)k0 * m + i;
#pragma unroll 4
        for (int t = g; t < kn; t += 4) {
          const T gv = gcol[(int64_t)t * m];
#pragma unroll
          for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] += gv * zs[t][c];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) red[g][lane][c] = acc[c];
    __syncthreads();
    if (split_combine<T>(red, nchunk, chunk, slot0, pc, part, counter, cidx) && rin) {
      for (int c = g; c < pc; c += 4) {
        const T v = red[0][lane][c];
        if (i < ncol) Y[first + i + (int64_t)(c0 + c) * n] = v;
        else U[S.sn_rowoff[s] + i + (int64_t)(c0 + c) * ldu] += v;
      }
    }
    __syncthreads();
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

// x1 = G^T [y1; x_cb] for a tile of 64 pivot columns of a front and one chunk of its rows
// i >= the tile's first column (G11 lower triangular): y1 from Y (pivot order), x_cb from X at the
// front's contribution rows (rows of ancestors, final already). The result is scattered into X at
// the pivots' rows. 256 threads = 64 columns x 4 row groups; Gt read along rows.
// grid (column tiles, fronts, row chunks).
template <class T, class R>
__global__ void __launch_bounds__(256) k_bwd_gemv(DevSym S, DevNum<T> N, const T* __restrict__ Gt,
                                                   const int32_t* __restrict__ list, int p, const T* __restrict__ Y,
                                                   R* X, int64_t ldx, T* __restrict__ part, int* __restrict__ counter) {
  __shared__ T ws[kGvChunk][kMaxRhsChunk];
  __shared__ T red[4][kGvTile][kMaxRhsChunk + 1];
  const int f = blockIdx.y;
  const int s = list[f];
  const int m = S.sn_m[s], ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int tile = blockIdx.x, chunk = blockIdx.z;
  const int j0 = tile * kGvTile;
  if (j0 >= ncol) return;
  const int nchunk = (m - j0 + kGvSplit - 1) / kGvSplit;
  if (chunk >= nchunk) return;
  const int ibeg = j0 + chunk * kGvSplit, istop = min(m, ibeg + kGvSplit);
  const int n = S.n;
  const int64_t roff = S.sn_rowoff[s];
  const int32_t* rows = S.rows_orig + roff;
  const T* Gs = Gt + S.panel_off[s];
  const int lane = threadIdx.x % kGvTile, g = threadIdx.x / kGvTile;
  const int j = j0 + lane;
  const bool jin = j < ncol;
  const int nrc = (p + kMaxRhsChunk - 1) / kMaxRhsChunk;
  for (int c0 = 0; c0 < p; c0 += kMaxRhsChunk) {
    const int pc = min(kMaxRhsChunk, p - c0);
    const int cidx = (f * (int)gridDim.x + tile) * nrc + c0 / kMaxRhsChunk;
    const int64_t slot0 = (int64_t)cidx * gridDim.z;
    T acc[kMaxRhsChunk];
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = (T)0;
    for (int i0 = ibeg; i0 < istop; i0 += kGvChunk) {
      const int kn = min(kGvChunk, istop - i0);
      __syncthreads();
      for (int idx = threadIdx.x; idx < kGvChunk * kMaxRhsChunk; idx += 256) {
        const int t = idx % kGvChunk, c = idx / kGvChunk;
        T v = (T)0;
        if (t < kn && c < pc) {
          const int i = i0 + t;
          v = i < ncol ? Y[first + i + (int64_t)(c0 + c) * n] : (T)X[rows[i] + (int64_t)(c0 + c) * ldx];
        }
        ws[t][c] = v;
      }
      __syncthreads();
      if (jin) {
        const T* grow = Gs + (int64_t)i0 * ncol + j;
#pragma unroll 4
        for (int t = g; t < kn; t += 4) {
          const T gv = grow[(int64_t)t * ncol];
#pragma unroll
          for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] += gv * ws[t][c];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) red[g][lane][c] = acc[c];
    __syncthreads();
    if (split_combine<T>(red, nchunk, chunk, slot0, pc, part, counter, cidx) && jin) {
      const int64_t xrow = rows[N.piv[first + j]];
      for (int c = g; c < pc; c += 4) X[xrow + (int64_t)(c0 + c) * ldx] = (R)red[0][lane][c];
    }
    __syncthreads();
  }
}

}  // namespace hsg
