// Multifrontal forward / backward substitution with p right-hand sides in partitioned-inverse
// form (the solve phase the reference hands to MUMPS, JOB = 3,
// src/factor/mumps_factorization.cpp:458-527).
//
// With P K P^T = L D L^T and, per front s (ncol pivots, m rows), L = [L11; L21], the solve phase
// reads, per front,
//     G  = [ G11 ; G21 ] = [ L11^-1 ; -L21 L11^-1 ]   (m x ncol, column-major, the panel layout)
//     Gt = G stored row-major                         (same offsets, element (i, j) at i*ncol + j)
// because L^-1 restricted to one front is [L11^-1 0; -L21 L11^-1 I]. Then
//   forward, bottom-up per level:  one GEMV per front, [ y1 ; u ] = G z1 + [ 0 ; z_cb ], where z =
//                                  own right-hand-side rows + children's update vectors, summed
//                                  in a fixed order (pull map) where z1 is staged and where the
//                                  contribution rows are written
//   diagonal:                      y = D^-1 y (1x1 and 2x2 blocks)
//   backward, top-down per level:  x1 = G^T [ y1 ; x_cb ] with x_cb the solution at the
//                                  front's contribution rows (rows of ancestors, final).
// One launch per level forward and one backward (each with a second one where long fronts split
// their products), and no sequential chain inside a front.
// Inside a solve the right-hand sides live in pivot numbering with the p columns interleaved
// (Xp[pos * p + c], the same for Y and the update rows U[(front row) * p + c]): a front's rows are
// contiguous and a gathered row is one contiguous run of p values. The caller's column-major X is
// permuted in at the start and out at the end (pos: pivot position of each original row).
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

// The children's update rows mapped onto one front row (the forward products' pull map): the
// first two inline (U row indices, -1 when absent), the others at cidx[b, e) in child order.
struct alignas(16) PullRow {
  int32_t u0, u1, b, e;
};

// A pointer into global memory as the compiler's address space 1: a value selected from two such
// pointers is still known to be global, so loads through it are global loads (a generic pointer
// gives flat loads, which the LDS wait counter also tracks).
template <class T>
using GlobalPtr = __attribute__((address_space(1))) T*;

// The sweeps stage right-hand-side rows with two threads per row, four values each: 128 rows per
// 256-thread pass.
static_assert(kGvChunk * 2 == 256 && kMaxRhsChunk == 8, "staging layout: 2 threads x 4 values per row");

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

// Pivot position of every original row: pos[rows_orig[rowoff + r]] = first + ipiv[first + r] for
// the pivot rows r < ncol of each front (one workgroup per front, all fronts).
// With them, per pivot position q, the pull map entry of the front row pivoted to q (the
// children's update rows summed into it by the forward products).
template <class T>
__global__ void __launch_bounds__(256) k_pivot_positions(DevSym S, DevNum<T> N, int32_t* __restrict__ pos,
                                                         const PullRow* __restrict__ cpull, PullRow* __restrict__ zpull) {
  const int s = blockIdx.x;
  const int ncol = S.sn_ncol[s], first = S.sn_first[s];
  const int64_t roff = S.sn_rowoff[s];
  for (int r = threadIdx.x; r < ncol; r += blockDim.x) {
    const int q = first + N.ipiv[first + r];
    pos[S.rows_orig[roff + r]] = q;
    zpull[q] = cpull[roff + r];
  }
}

// Pivot position of every front row (index rowoff + i): rpos = pos[rows_orig].
static __global__ void k_row_positions(int64_t rows_total, const int32_t* __restrict__ rows_orig,
                                const int32_t* __restrict__ pos, int32_t* __restrict__ rpos) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < rows_total; k += (int64_t)gridDim.x * blockDim.x)
    rpos[k] = pos[rows_orig[k]];
}

// Xp[pos[i] * p + c] = X[i + c * ldx] (into pivot numbering, columns interleaved).
template <class R, class T>
__global__ void k_permute_in(int n, int p, const R* __restrict__ X, int64_t ldx, const int32_t* __restrict__ pos,
                             T* __restrict__ Xp) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T* dst = Xp + (int64_t)pos[i] * p;
    for (int c = 0; c < p; ++c) dst[c] = (T)X[i + c * ldx];
  }
}

// X[i + c * ldx] = Xp[pos[i] * p + c] (back to the caller's numbering and layout).
template <class T, class R>
__global__ void k_permute_out(int n, int p, const T* __restrict__ Xp, const int32_t* __restrict__ pos, R* __restrict__ X,
                              int64_t ldx) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T* src = Xp + (int64_t)pos[i] * p;
    for (int c = 0; c < p; ++c) X[i + c * ldx] = (R)src[c];
  }
}

// Split products: a tile's long dimension (columns of G forward, rows of Gt backward) is cut into
// chunks, one workgroup each. With more than one chunk, each workgroup stores its partial sums
// (64 x 8 per group of right-hand sides; forward [value][row], backward [column][value]) in its
// slot of `part`, and a later launch adds the chunks' partials in chunk order: forward a second
// launch of the level, backward the next level's launch (its own tasks for the rows they read,
// extra workgroups for Xp). The sum does not depend on the order the workgroups ran in, and
// stream order makes the partials visible (no cross-workgroup synchronization inside a launch).
// Slot index: (slot of chunk) * rhs_groups + rhs_group.
constexpr int kGvSplit = 256;          // columns per chunk of a split forward product, at most
constexpr int kFwdTargetGroups = 1536;  // workgroups a forward level's grid should reach (6 per CU)
constexpr int kPartSlot = kGvTile * kMaxRhsChunk;  // values per partial slot

// src[0] + src[stride] + ... (n terms), added in that order. The loads are issued 32 at a time and
// without conditions (terms past the end read the last one and are not added), so that a block's
// loads are in flight together.
template <class T>
__device__ __forceinline__ T sum_partials(const T* __restrict__ src, int64_t stride, int n) {
  constexpr int kB = 32;
  T v = (T)0;
  for (int c0 = 0; c0 < n; c0 += kB) {
    T x[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) x[u] = src[(int64_t)min(c0 + u, n - 1) * stride];
#pragma unroll
    for (int u = 0; u < kB; ++u)
      if (c0 + u < n) v += x[u];
  }
  return v;
}

// One task of the forward sweep: rows [r0, r0 + 64) of front s's G and its columns [kbeg, kstop)
// (a chunk of the columns up to min(ncol, r0 + 64): G11 is lower triangular). Everything a task
// needs about its front is stored in the task, so the first memory accesses of a workgroup are
// the G entries and the right-hand-side gathers themselves.
struct FwdTask {
  int32_t m, ncol, first, r0;   // front rows, pivots, first pivot position; first row of the tile
  int32_t kbeg, kstop, nchunk;  // columns of this chunk; chunks of the tile (> 1: split tile)
  int32_t pad;
  int64_t goff;                 // offset of the front's panel in G (column-major, ld = m)
  int64_t rowoff;               // offset of the front's rows in the row lists (U, the pull map)
  int64_t slot;                 // partial slot of this chunk (split tiles; chunk 0's in a head), else -1
};

constexpr int kFwdCols = kGvChunk / 4;  // G entries per thread and pass at most (4 column groups)

// [y1; u] = G z1 + [0; u_children] over one task: rows < ncol (G11) give y1 in Y, contribution
// rows get the children's update rows plus G21 z1 in U; a split tile's chunks store partial sums
// instead (k_fwd_reduce adds them). 256 threads = 64 rows x 4 column groups; per pass of up to
// kGvChunk columns a thread's C G entries (columns g, g + 4, ...; C >= the columns per thread of
// every pass of the launch, columns past the pass's end read its last column) are loaded into
// registers before z is staged, so the G loads, the gathers and (for contribution rows) the
// children's sums of the epilogue are in flight together. Products and sums run in the same order
// as one pass over the columns: results do not depend on the pass structure.
template <class T, int C>
__global__ void __launch_bounds__(256) k_fwd_task(const FwdTask* __restrict__ tasks, int p, const T* __restrict__ G,
                                                   const T* __restrict__ Xp, T* __restrict__ Y, T* __restrict__ U,
                                                   T* __restrict__ part, const PullRow* __restrict__ zpull,
                                                   const PullRow* __restrict__ cpull, const int32_t* __restrict__ cidx) {
  __shared__ T zs[kGvChunk][kMaxRhsChunk];
  __shared__ T red[4][kGvTile][kMaxRhsChunk + 1];
  const FwdTask tk = tasks[blockIdx.x];
  const int lane = threadIdx.x % kGvTile;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / kGvTile);
  const int i = tk.r0 + lane;
  const bool rin = i < tk.m;
  const T* grow = G + tk.goff + (rin ? i : tk.r0);
  const int64_t r = tk.rowoff + (rin ? i : tk.r0);
  const bool cb_out = tk.nchunk == 1 && rin && i >= tk.ncol;  // writes a contribution row
  const int nrc = (p + kMaxRhsChunk - 1) / kMaxRhsChunk;
  // staging layout: two threads per row of z, four values each
  const int st = threadIdx.x >> 1, sh = (threadIdx.x & 1) * 4;
  // the children's rows mapped onto row r (loaded whatever the row, used where it writes a
  // contribution row: a load under a condition would be waited for at once)
  const PullRow er = cpull[r];
  // one pass over columns [k0, k0 + kGvChunk) of the chunk for one group of right-hand sides
  auto pass = [&](int c0, int pc, int k0, T(&acc)[kMaxRhsChunk], T(&csum)[2], bool first_pass)
                  __attribute__((always_inline)) {
    const int kn = min(kGvChunk, tk.kstop - k0);
    // z row st of the pass: pull map entry and right-hand side first (they need the row index only)
    const bool sval = st < kn;
    const int64_t zq = (int64_t)tk.first + k0 + (sval ? st : 0);
    const PullRow zr = zpull[zq];
    T zv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) zv[c] = Xp[zq * p + c0 + min(sh + c, pc - 1)];
    // this thread's G entries of the pass (columns g, g + 4, ...)
    T gv[C];
#pragma unroll
    for (int q = 0; q < C; ++q) gv[q] = grow[(int64_t)(k0 + min(g + 4 * q, kn - 1)) * tk.m];
    // the children's update rows, in child order: of z row st, and (first pass) of the
    // contribution row r. The two inline ones of each are loaded whatever the row (row 0 stands in
    // for an absent one) before any of them is used; the others (rare) after.
    T za[4], zb[4], ea[2], eb[2];
    {
      const T* ua = U + (int64_t)max(zr.u0, 0) * p + c0;
      const T* ub = U + (int64_t)max(zr.u1, 0) * p + c0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        za[c] = ua[min(sh + c, pc - 1)];
        zb[c] = ub[min(sh + c, pc - 1)];
      }
    }
    const int c1 = min(g, pc - 1), c2 = min(g + 4, pc - 1);
    if (first_pass) {
      const T* ua = U + (int64_t)max(er.u0, 0) * p + c0;
      const T* ub = U + (int64_t)max(er.u1, 0) * p + c0;
      ea[0] = ua[c1];
      ea[1] = ua[c2];
      eb[0] = ub[c1];
      eb[1] = ub[c2];
    }
    {
      const bool ha = sval && zr.u0 >= 0, hb = sval && zr.u1 >= 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        zv[c] = ha ? zv[c] + za[c] : zv[c];
        zv[c] = hb ? zv[c] + zb[c] : zv[c];
      }
      for (int q = zr.b; sval && q < zr.e; ++q) {
        const T* u = U + (int64_t)cidx[q] * p + c0;
#pragma unroll
        for (int c = 0; c < 4; ++c) zv[c] += u[min(sh + c, pc - 1)];
      }
    }
    if (first_pass) {
      // csum starts at +0 and never becomes -0, so adding +0 for an absent row changes nothing (and
      // keeps the loads unconditional)
      const bool ha = cb_out && er.u0 >= 0, hb = cb_out && er.u1 >= 0;
      csum[0] += ha ? ea[0] : (T)0;
      csum[1] += ha ? ea[1] : (T)0;
      csum[0] += hb ? eb[0] : (T)0;
      csum[1] += hb ? eb[1] : (T)0;
      for (int q = er.b; cb_out && q < er.e; ++q) {
        const T* u = U + (int64_t)cidx[q] * p + c0;
        csum[0] += u[c1];
        csum[1] += u[c2];
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) zs[st][sh + c] = (sval && sh + c < pc) ? zv[c] : (T)0;
    __syncthreads();
    if (rin) {
#pragma unroll
      for (int q = 0; q < C; ++q)
        if (g + 4 * q < kn) {
#pragma unroll
          for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] += gv[q] * zs[g + 4 * q][c];
        }
    }
  };
  // one group of at most kMaxRhsChunk right-hand sides; its first pass (and the first group) is not
  // inside a loop, so that the gathers need not wait for the G loads still in flight
  auto rhs_group = [&](int c0) __attribute__((always_inline)) {
    const int pc = min(kMaxRhsChunk, p - c0);
    T acc[kMaxRhsChunk];
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = (T)0;
    T csum[2] = {(T)0, (T)0};  // children's update rows at values c0 + g, c0 + g + 4 of row r
    pass(c0, pc, tk.kbeg, acc, csum, true);
    for (int k0 = tk.kbeg + kGvChunk; k0 < tk.kstop; k0 += kGvChunk) pass(c0, pc, k0, acc, csum, false);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) red[g][lane][c] = acc[c];
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int c = g + 4 * h;
      if (c >= pc) continue;
      const T v = red[0][lane][c] + red[1][lane][c] + red[2][lane][c] + red[3][lane][c];
      if (tk.nchunk > 1) {
        part[((tk.slot * nrc + c0 / kMaxRhsChunk) * kPartSlot) + c * kGvTile + lane] = v;
      } else if (rin) {
        if (i < tk.ncol) Y[(int64_t)(tk.first + i) * p + c0 + c] = v;
        else U[r * p + c0 + c] = csum[h] + v;
      }
    }
    __syncthreads();
  };
  rhs_group(0);
  for (int c0 = kMaxRhsChunk; c0 < p; c0 += kMaxRhsChunk) rhs_group(c0);
}

// Forward tasks of the levels whose tiles have at most kWaveCols columns in one chunk: one task
// per 64-lane wave, four per workgroup (a 256-thread workgroup per small tile leaves most of its
// threads without work and holds the device's wave slots, which concurrent solves share). Lane =
// row of the tile: a lane loads its row's C entries of G, the wave stages the tile's z rows
// (right-hand side + children's update rows) into an LDS area of its own, and every lane forms its
// row of [y1; u] alone, columns in order, so no reduction between waves is needed. The pull-map
// and right-hand-side loads are issued before the G loads, all children's rows before their use.
constexpr int kWaveCols = 32;

template <class T, int C>
__global__ void __launch_bounds__(256) k_fwd_wave(const FwdTask* __restrict__ tasks, int ntasks, int p,
                                                   const T* __restrict__ G, const T* __restrict__ Xp,
                                                   T* __restrict__ Y, T* __restrict__ U,
                                                   const PullRow* __restrict__ zpull, const PullRow* __restrict__ cpull,
                                                   const int32_t* __restrict__ cidx) {
  static_assert(C <= kWaveCols && 2 * kWaveCols == 64, "staging: 32 rows x 2 halves per wave");
  __shared__ T zs[4][kWaveCols][kMaxRhsChunk];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int lane = threadIdx.x % 64;
  const int ti = blockIdx.x * 4 + w;
  const bool live = ti < ntasks;
  const FwdTask tk = tasks[live ? ti : ntasks - 1];
  const int kn = tk.kstop - tk.kbeg;
  const int i = tk.r0 + lane;
  const bool rin = live && i < tk.m;
  const T* grow = G + tk.goff + (rin ? i : tk.r0);
  const int64_t r = tk.rowoff + (rin ? i : tk.r0);
  const bool cb_out = rin && i >= tk.ncol;
  const int st = lane >> 1, sh = (lane & 1) * 4;  // staging: two lanes per z row, four values each
  const bool sval = live && st < kn;
  const int64_t zq = (int64_t)tk.first + tk.kbeg + (sval ? st : 0);
  auto rhs_group = [&](int c0) __attribute__((always_inline)) {
    const int pc = min(kMaxRhsChunk, p - c0);
    // pull-map entries and right-hand side first (they need the row indices only)
    const PullRow zr = zpull[zq];
    const PullRow er = cpull[r];
    T zv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) zv[c] = Xp[zq * p + c0 + min(sh + c, pc - 1)];
    T gv[C];
#pragma unroll
    for (int q = 0; q < C; ++q) gv[q] = grow[(int64_t)(tk.kbeg + min(q, kn - 1)) * tk.m];
    // the inline children's rows of z row st and of row r (row 0 stands in for an absent one)
    T za[4], zb[4], ea[kMaxRhsChunk], eb[kMaxRhsChunk];
    {
      const T* ua = U + (int64_t)max(zr.u0, 0) * p + c0;
      const T* ub = U + (int64_t)max(zr.u1, 0) * p + c0;
      const T* va = U + (int64_t)max(er.u0, 0) * p + c0;
      const T* vb = U + (int64_t)max(er.u1, 0) * p + c0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        za[c] = ua[min(sh + c, pc - 1)];
        zb[c] = ub[min(sh + c, pc - 1)];
      }
#pragma unroll
      for (int c = 0; c < kMaxRhsChunk; ++c) {
        ea[c] = va[min(c, pc - 1)];
        eb[c] = vb[min(c, pc - 1)];
      }
    }
    {
      const bool ha = sval && zr.u0 >= 0, hb = sval && zr.u1 >= 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        zv[c] = ha ? zv[c] + za[c] : zv[c];
        zv[c] = hb ? zv[c] + zb[c] : zv[c];
      }
      for (int q = zr.b; sval && q < zr.e; ++q) {
        const T* u = U + (int64_t)cidx[q] * p + c0;
#pragma unroll
        for (int c = 0; c < 4; ++c) zv[c] += u[min(sh + c, pc - 1)];
      }
    }
    // the children's sum of contribution row r: starts at +0 and never becomes -0, so adding +0 for
    // an absent row changes nothing
    T csum[kMaxRhsChunk];
    {
      const bool ha = cb_out && er.u0 >= 0, hb = cb_out && er.u1 >= 0;
#pragma unroll
      for (int c = 0; c < kMaxRhsChunk; ++c) csum[c] = ((T)0 + (ha ? ea[c] : (T)0)) + (hb ? eb[c] : (T)0);
      for (int q = er.b; cb_out && q < er.e; ++q) {
        const T* u = U + (int64_t)cidx[q] * p + c0;
#pragma unroll
        for (int c = 0; c < kMaxRhsChunk; ++c) csum[c] += u[min(c, pc - 1)];
      }
    }
    __syncthreads();  // the previous group of right-hand sides is done with the staging area
#pragma unroll
    for (int c = 0; c < 4; ++c) zs[w][st][sh + c] = (sval && sh + c < pc) ? zv[c] : (T)0;
    __syncthreads();
    T acc[kMaxRhsChunk];
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = (T)0;
#pragma unroll
    for (int q = 0; q < C; ++q)
      if (q < kn) {
#pragma unroll
        for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] += gv[q] * zs[w][q][c];
      }
    if (rin) {
      for (int c = 0; c < pc; ++c) {
        if (i < tk.ncol) Y[(int64_t)(tk.first + i) * p + c0 + c] = acc[c];
        else U[r * p + c0 + c] = csum[c] + acc[c];
      }
    }
  };
  rhs_group(0);
  for (int c0 = kMaxRhsChunk; c0 < p; c0 += kMaxRhsChunk) rhs_group(c0);
}

// Second launch of a split forward level: one workgroup per split row tile (its chunk-0 task in
// `heads`), the chunks' partials added in chunk order, written like k_fwd_task's result.
template <class T>
__global__ void __launch_bounds__(256) k_fwd_reduce(const FwdTask* __restrict__ heads, int p, T* __restrict__ Y,
                                                     T* __restrict__ U, const T* __restrict__ part,
                                                     const PullRow* __restrict__ cpull, const int32_t* __restrict__ cidx) {
  // 64 rows x 4 value groups; a thread adds the partials of values g, g + 4 of its row (all their
  // loads issued together) and, for a contribution row, the children's rows (inline ones loaded
  // whatever the row, as in the products)
  const FwdTask tk = heads[blockIdx.x];
  const int nrc = (p + kMaxRhsChunk - 1) / kMaxRhsChunk;
  const int lane = threadIdx.x % kGvTile, g = threadIdx.x / kGvTile;
  const int i = tk.r0 + lane;
  const bool rin = i < tk.m;
  const int64_t r = tk.rowoff + (rin ? i : tk.r0);
  const bool cb_out = rin && i >= tk.ncol;
  const PullRow er = cpull[r];
  for (int c0 = 0; c0 < p; c0 += kMaxRhsChunk) {
    const int pc = min(kMaxRhsChunk, p - c0);
    const int c1 = min(g, pc - 1), c2 = min(g + 4, pc - 1);
    const T* src = part + (tk.slot * nrc + c0 / kMaxRhsChunk) * (int64_t)kPartSlot + lane;
    const T v1 = sum_partials<T>(src + c1 * kGvTile, (int64_t)nrc * kPartSlot, tk.nchunk);
    const T v2 = sum_partials<T>(src + c2 * kGvTile, (int64_t)nrc * kPartSlot, tk.nchunk);
    const T* ua = U + (int64_t)max(er.u0, 0) * p + c0;
    const T* ub = U + (int64_t)max(er.u1, 0) * p + c0;
    const T a1 = ua[c1], a2 = ua[c2], b1 = ub[c1], b2 = ub[c2];
    // the children's sum starts at +0 and never becomes -0: adding +0 for an absent row changes
    // nothing
    const bool ha = cb_out && er.u0 >= 0, hb = cb_out && er.u1 >= 0;
    T s1 = ((T)0 + (ha ? a1 : (T)0)) + (hb ? b1 : (T)0);
    T s2 = ((T)0 + (ha ? a2 : (T)0)) + (hb ? b2 : (T)0);
    for (int q = er.b; cb_out && q < er.e; ++q) {
      const T* u = U + (int64_t)cidx[q] * p + c0;
      s1 += u[c1];
      s2 += u[c2];
    }
    if (rin) {
      if (g < pc) {
        if (i < tk.ncol) Y[(int64_t)(tk.first + i) * p + c0 + g] = v1;
        else U[r * p + c0 + g] = s1 + v1;
      }
      if (g + 4 < pc) {
        if (i < tk.ncol) Y[(int64_t)(tk.first + i) * p + c0 + g + 4] = v2;
        else U[r * p + c0 + g + 4] = s2 + v2;
      }
    }
  }
}

// ---------------------------------------------------------------------------------- diagonal

// Y (pivot numbering, p columns interleaved) = D^-1 Y.
template <class T>
__global__ void k_solve_diag(int n, const T* __restrict__ dval, const T* __restrict__ doff,
                             const int8_t* __restrict__ ptype, T* Y, int p) {
  const int64_t total = (int64_t)n * p;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / p);
    const int8_t t = ptype[i];
    T* z = Y + idx;
    if (t == kPiv1) {
      z[0] = z[0] / dval[i];
    } else if (t == kPiv2First) {
      const T a = dval[i], bb = doff[i], d = dval[i + 1];
      const T det = a * d - bb * bb;
      const T z1 = z[0], z2 = z[p];
      z[0] = (d * z1 - bb * z2) / det;
      z[p] = (a * z2 - bb * z1) / det;
    } else if (t == kPivNull) {
      z[0] = (T)0;
    }
  }
}

// ---------------------------------------------------------------------------------- backward

// One task of the backward sweep: columns [j0, j0 + jn) of front s (jw = the power of two >= jn
// that the 256 threads are laid out by: jw columns x 256 / jw row groups) and rows [i0, i0 + in)
// of its Gt (in <= kBwdRows), chunk `chunk` of the nchunk row chunks of that column tile (slot0:
// the slot of its chunk 0 when nchunk > 1). The front's data the task needs is stored in it.
struct BwdTask {
  int32_t s, j0, jn, jw, i0, in, chunk, nchunk;
  int64_t slot0;
  int32_t ncol, first;  // the front's pivots and first pivot position
  int64_t goff;         // Gt offset of the task's element (i0, j0)
  int64_t rpoff;        // offset of row i0 of the front in the row lists (pivot positions rpos)
};

constexpr int kBwdRows = 512;  // rows of a backward task at most (staged in LDS)

// Per pivot position q: where its backward value stands while the level that computes it ends. A
// pivot of a split column tile has partial sums instead (slot0: the tile's chunk-0 slot, nchunk
// chunks, jj: its column in the tile) until the next level's launch, which adds them; level = the
// height of its front. slot0 = -1 for the pivots of unsplit tiles.
struct alignas(16) PivSplit {
  int32_t slot0, nchunk, jj, level;
};

// x1 = G^T [y1; x_cb] over one task: y1 from Y (pivot rows), x_cb from Xp at the pivot positions
// of the front's contribution rows (rows of ancestors, final already); the result goes to Xp at
// the front's own pivot positions, or, for a split tile, to the task's partial slot. Gt is read
// along rows (jw consecutive columns per row group). A thread's R rows (grp, grp + ngrp, ...) of
// Gt are loaded into registers first, so they are in flight while the rows of [y1; x_cb] are
// gathered (pivot positions, then values) and staged; the products run in row order as in one
// pass over the rows. R >= the rows per thread of every task of the launch.
template <class T, int R>
__global__ void __launch_bounds__(256) k_bwd_task(const T* __restrict__ Gt, const BwdTask* __restrict__ tasks, int ntasks,
                                                   const BwdTask* __restrict__ heads, int p, const T* __restrict__ Y,
                                                   T* Xp, const int32_t* __restrict__ rpos, const PivSplit* __restrict__ psplit,
                                                   int level, const T* __restrict__ part_in, T* __restrict__ part) {
  static_assert(R >= 1 && R <= kBwdRows / 8, "rows per thread");
  __shared__ T ws[kBwdRows][kMaxRhsChunk];
  T(*red)[kMaxRhsChunk + 1] = reinterpret_cast<T(*)[kMaxRhsChunk + 1]>(&ws[0][0]);  // after the products
  static_assert(256 * (kMaxRhsChunk + 1) <= kBwdRows * kMaxRhsChunk, "reduction area inside the staging area");
  const int nrc = (p + kMaxRhsChunk - 1) / kMaxRhsChunk;
  if ((int)blockIdx.x >= ntasks) {
    // the workgroups past the tasks: the split tiles of the level above, their partials (part_in)
    // added in chunk order and written to Xp for the levels below this one
    const BwdTask hd = heads[blockIdx.x - ntasks];
    for (int idx = threadIdx.x; idx < hd.jn * p; idx += blockDim.x) {
      const int j2 = idx % hd.jn, c = idx / hd.jn;
      const int q = c / kMaxRhsChunk, cc = c % kMaxRhsChunk;
      const T v = sum_partials<T>(part_in + (hd.slot0 * nrc + q) * (int64_t)kPartSlot + j2 * kMaxRhsChunk + cc,
                                  (int64_t)nrc * kPartSlot, hd.nchunk);
      Xp[(int64_t)(hd.first + hd.j0 + j2) * p + c] = v;
    }
    return;
  }
  const BwdTask tk = tasks[blockIdx.x];
  const int JW = tk.jw, ngrp = 256 / JW;
  const int tid = threadIdx.x, jj = tid % JW, grp = tid / JW;
  const bool jin = jj < tk.jn;
  const int st = tid >> 1, sh = (tid & 1) * 4;  // staging: two threads per row, four values each
  constexpr int kRounds = kBwdRows / 128;
  // pivot positions of the staged rows first (the value loads wait for them, and the memory
  // counter is in order), then the R rows of Gt: all R loads are issued (rows past the task's end
  // read its last row), since loads under a condition would be waited for one by one
  const int32_t* rp = rpos + tk.rpoff;
  int32_t pos[kRounds];
#pragma unroll
  for (int k = 0; k < kRounds; ++k) pos[k] = rp[min(st + 128 * k, tk.in - 1)];
  const T* g0 = Gt + tk.goff + (jin ? jj : 0);
  T gv[R];
#pragma unroll
  for (int q = 0; q < R; ++q) gv[q] = g0[(int64_t)min(grp + q * ngrp, tk.in - 1) * tk.ncol];
  // one group of at most kMaxRhsChunk right-hand sides; the first group is not inside a loop, so
  // that its value loads need not wait for the Gt loads still in flight
  auto rhs_group = [&](int c0) __attribute__((always_inline)) {
    const int pc = min(kMaxRhsChunk, p - c0);
    // rows st + 128 k of [y1; x_cb] (rows past the task's end read its last row; global pointers, so
    // that the loads are not counted with the LDS accesses), with the split state of the
    // contribution rows' pivot positions
    T v[kRounds][4];
    PivSplit ps[kRounds];
#pragma unroll
    for (int k = 0; k < kRounds; ++k) {
      if (k * 128 >= tk.in) break;
      const int i = tk.i0 + min(st + 128 * k, tk.in - 1);
      const GlobalPtr<const T> src = i < tk.ncol ? GlobalPtr<const T>(Y) + (int64_t)(tk.first + i) * p + c0
                                                 : GlobalPtr<const T>(Xp) + (int64_t)pos[k] * p + c0;
      ps[k] = psplit[i < tk.ncol ? tk.first + i : pos[k]];
#pragma unroll
      for (int c = 0; c < 4; ++c) v[k][c] = src[min(sh + c, pc - 1)];
    }
    __syncthreads();  // the previous group of right-hand sides is done with the staging area
#pragma unroll
    for (int k = 0; k < kRounds; ++k) {
      if (k * 128 >= tk.in) break;
      const bool valid = st + 128 * k < tk.in;
#pragma unroll
      for (int c = 0; c < 4; ++c) ws[st + 128 * k][sh + c] = (valid && sh + c < pc) ? v[k][c] : (T)0;
    }
    // a contribution row on a split tile of the level just above: its value is still in that
    // level's partials (this launch's extra workgroups write it to Xp for the levels below), added
    // here in chunk order as they do, over the staged value
#pragma unroll
    for (int k = 0; k < kRounds; ++k) {
      if (k * 128 >= tk.in) break;
      const int t = st + 128 * k;
      if (t < tk.in && tk.i0 + t >= tk.ncol && ps[k].level == level + 1 && ps[k].slot0 >= 0) {
        // the row's four values are contiguous in every chunk's slot: eight chunks' loads at a time
        const T* src = part_in + (ps[k].slot0 * (int64_t)nrc + c0 / kMaxRhsChunk) * kPartSlot +
                       ps[k].jj * kMaxRhsChunk + sh;
        const int64_t stride = (int64_t)nrc * kPartSlot;
        const int n = ps[k].nchunk;
        T a[4] = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll 1
        for (int k0 = 0; k0 < n; k0 += 8) {
          T x[8][4];
#pragma unroll
          for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int c = 0; c < 4; ++c) x[u][c] = src[(int64_t)min(k0 + u, n - 1) * stride + c];
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (k0 + u < n)
#pragma unroll
              for (int c = 0; c < 4; ++c) a[c] += x[u][c];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (sh + c < pc) ws[t][sh + c] = a[c];
      }
    }
    __syncthreads();
    T acc[kMaxRhsChunk];
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] = (T)0;
    if (jin) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int t = grp + q * ngrp;
        if (t < tk.in) {
#pragma unroll
          for (int c = 0; c < kMaxRhsChunk; ++c) acc[c] += gv[q] * ws[t][c];
        }
      }
    }
    __syncthreads();  // every thread is done reading the staged rows
#pragma unroll
    for (int c = 0; c < kMaxRhsChunk; ++c) red[tid][c] = acc[c];
    __syncthreads();
    T* slot = tk.nchunk > 1 ? part + ((tk.slot0 + tk.chunk) * nrc + c0 / kMaxRhsChunk) * (int64_t)kPartSlot : nullptr;
    for (int idx = tid; idx < JW * pc; idx += 256) {
      const int j2 = idx % JW, c = idx / JW;
      if (j2 >= tk.jn) continue;
      T s = (T)0;
      for (int g2 = 0; g2 < ngrp; ++g2) s += red[g2 * JW + j2][c];
      if (slot) slot[j2 * kMaxRhsChunk + c] = s;
      else Xp[(int64_t)(tk.first + tk.j0 + j2) * p + c0 + c] = s;
    }
  };
  rhs_group(0);
  for (int c0 = kMaxRhsChunk; c0 < p; c0 += kMaxRhsChunk) rhs_group(c0);
}

// Second launch of a split backward level: one workgroup per split column tile (the task of its
// chunk 0, in `heads`), the chunks' partials added in chunk order and written to Xp.
template <class T>
__global__ void __launch_bounds__(256) k_bwd_reduce(const BwdTask* __restrict__ heads, int p, T* Xp,
                                                     const T* __restrict__ part) {
  const BwdTask tk = heads[blockIdx.x];
  const int nrc = (p + kMaxRhsChunk - 1) / kMaxRhsChunk;
  for (int idx = threadIdx.x; idx < tk.jn * p; idx += blockDim.x) {
    const int j2 = idx % tk.jn, c = idx / tk.jn;
    const int q = c / kMaxRhsChunk, cc = c % kMaxRhsChunk;
    const T v = sum_partials<T>(part + (tk.slot0 * nrc + q) * (int64_t)kPartSlot + j2 * kMaxRhsChunk + cc,
                                (int64_t)nrc * kPartSlot, tk.nchunk);
    Xp[(int64_t)(tk.first + tk.j0 + j2) * p + c] = v;
  }
}

}  // namespace hsg
