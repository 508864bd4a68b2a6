// Sparse and dense kernels of the shift-invert operator and the block Krylov-Schur solver.
//
// All HBM-bound: every kernel streams its n-long operands once. Replaces the reference's CPU
// kernels spmm (src/sparse/kernels.cpp:797), shifted_residual (kernels.cpp:850), gram and the
// block CGS2 projections (src/linalg/dense_ops.cpp:752, :883) and rotate_in_place
// (src/eigen/block_lanczos.hpp:96-98). Reductions use a fixed two-pass tree (per-workgroup
// partials, then one ordered sum), so results are bitwise reproducible run to run.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>

#include "krylov.hpp"

namespace hsg {

namespace {

int grid_for(int64_t n, int per, int cap) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(cap, (n + per - 1) / per));
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

// Scalars of the kernels: double, or zval (complex, interleaved real and imaginary parts as in
// std::complex<double>). The real instantiations perform exactly the operations of plain double
// code (conj is the identity, |a| is fabs).
struct zval {
  double re, im;
};
__device__ __forceinline__ zval operator+(zval a, zval b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ zval operator-(zval a, zval b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ zval operator*(zval a, zval b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ zval operator*(double a, zval b) { return {a * b.re, a * b.im}; }
__device__ __forceinline__ zval& operator+=(zval& a, zval b) {
  a.re += b.re;
  a.im += b.im;
  return a;
}
__device__ __forceinline__ double conj_of(double a) { return a; }
__device__ __forceinline__ zval conj_of(zval a) { return {a.re, -a.im}; }
__device__ __forceinline__ double abs_of(double a) { return fabs(a); }
__device__ __forceinline__ double abs_of(zval a) { return hypot(a.re, a.im); }
__device__ __forceinline__ double abs2_of(double a) { return a * a; }
__device__ __forceinline__ double abs2_of(zval a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ bool finite_of(double a) { return isfinite(a); }
__device__ __forceinline__ bool finite_of(zval a) { return isfinite(a.re) && isfinite(a.im); }
__device__ __forceinline__ zval operator*(zval a, double b) { return {a.re * b, a.im * b}; }
template <class S>
__device__ __forceinline__ S zero_of() {
  return S{};
}
__device__ __forceinline__ double shfl_xor_of(double v, int off) { return __shfl_xor(v, off); }
__device__ __forceinline__ zval shfl_xor_of(zval v, int off) { return {__shfl_xor(v.re, off), __shfl_xor(v.im, off)}; }

// ------------------------------------------------------------------------------ sparse kernels

// Y = alpha M X + beta Y, one thread per row (rows of Yee pencils hold ~13 entries).
template <class S>
__global__ void k_csr_spmm(int64_t n, const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                           const S* __restrict__ v, const S* __restrict__ x, int64_t ldx, S* __restrict__ y,
                           int64_t ldy, int p, double alpha, double beta) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = rp[i], e = rp[i + 1];
    for (int c0 = 0; c0 < p; c0 += 8) {
      const int pc = min(8, p - c0);
      S acc[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[c] = zero_of<S>();
      for (int64_t q = b; q < e; ++q) {
        const S a = v[q];
        const int64_t j = ci[q];
#pragma unroll
        for (int c = 0; c < 8; ++c)
          if (c < pc) acc[c] += a * x[j + (c0 + c) * ldx];
      }
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (c < pc) {
          S* yy = y + i + (c0 + c) * ldy;
          *yy = beta == 0.0 ? alpha * acc[c] : alpha * acc[c] + beta * (*yy);
        }
    }
  }
}

// Y = alpha diag(d) X + beta Y (d real; NULL = identity).
template <class S>
__global__ void k_diag_spmm(int64_t n, const double* __restrict__ d, const S* __restrict__ x, int64_t ldx,
                            S* __restrict__ y, int64_t ldy, int p, double alpha, double beta) {
  const int64_t total = n * p;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx % n, c = idx / n;
    const double dv = d ? d[i] : 1.0;
    S* yy = y + i + c * ldy;
    const S t = (alpha * dv) * x[i + c * ldx];
    *yy = beta == 0.0 ? t : t + beta * (*yy);
  }
}

// R = B (X + sigma Y) - A Y with per-workgroup partial sums of |R|^2 and maxima of the
// componentwise quotient |R_i| / (|B| |W| + |A| |Y|)_i, W = X + sigma Y (0/0 = 0, r/0 = inf).
constexpr int kResThreads = 256;

template <class S>
__global__ void __launch_bounds__(kResThreads)
    k_shifted_residual(int64_t n, const int64_t* __restrict__ arp, const int32_t* __restrict__ aci,
                       const S* __restrict__ av, const int64_t* __restrict__ brp, const int32_t* __restrict__ bci,
                       const S* __restrict__ bv, const double* __restrict__ bd, double sigma, const S* __restrict__ x,
                       int64_t ldx, const S* __restrict__ y, int64_t ldy, S* __restrict__ r, int64_t ldr, int p,
                       double* __restrict__ part_sum, double* __restrict__ part_max) {
  __shared__ double ssum[kResThreads / 64][16];
  __shared__ double smax[kResThreads / 64][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c0 = 0; c0 < p; c0 += 16) {
    const int pc = min(16, p - c0);
    double lsum[16], lmax[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      lsum[c] = 0.0;
      lmax[c] = 0.0;
    }
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      for (int c = 0; c < pc; ++c) {
        const int64_t cc = c0 + c;
        S ay = zero_of<S>();
        double aay = 0.0;
        for (int64_t q = arp[i]; q < arp[i + 1]; ++q) {
          const S a = av[q], yy = y[aci[q] + cc * ldy];
          ay += a * yy;
          aay += abs_of(a) * abs_of(yy);
        }
        S bw = zero_of<S>();
        double bbw = 0.0;
        if (brp) {
          for (int64_t q = brp[i]; q < brp[i + 1]; ++q) {
            const int64_t j = bci[q];
            const S w = x[j + cc * ldx] + sigma * y[j + cc * ldy];
            bw += bv[q] * w;
            bbw += abs_of(bv[q]) * abs_of(w);
          }
        } else {
          const double dv = bd ? bd[i] : 1.0;
          const S w = x[i + cc * ldx] + sigma * y[i + cc * ldy];
          bw = dv * w;
          bbw = fabs(dv) * abs_of(w);
        }
        const S ri = bw - ay;
        r[i + cc * ldr] = ri;
        lsum[c] += abs2_of(ri);
        const double s = bbw + aay;
        const double ra = abs_of(ri);
        const double q = s > 0.0 ? ra / s : (ra == 0.0 ? 0.0 : INFINITY);
        lmax[c] = (finite_of(ri) && isfinite(s)) ? fmax(lmax[c], q) : INFINITY;
      }
    }
    for (int c = 0; c < pc; ++c) {
      double s = lsum[c], m = lmax[c];
      for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        m = fmax(m, __shfl_xor(m, off));
      }
      if (lane == 0) {
        ssum[wave][c] = s;
        smax[wave][c] = m;
      }
    }
    __syncthreads();
    if (threadIdx.x < pc) {
      double s = 0.0, m = 0.0;
      for (int w = 0; w < kResThreads / 64; ++w) {
        s += ssum[w][threadIdx.x];
        m = fmax(m, smax[w][threadIdx.x]);
      }
      part_sum[(int64_t)blockIdx.x * p + c0 + threadIdx.x] = s;
      part_max[(int64_t)blockIdx.x * p + c0 + threadIdx.x] = m;
    }
    __syncthreads();
  }
}

// Ordered reduction of per-workgroup partials: out[c] = sum (or max) over g of part[g * p + c].
__global__ void k_reduce_partials(int groups, int p, const double* __restrict__ part, double* __restrict__ out,
                                  int is_max) {
  const int c = blockIdx.x;
  __shared__ double red[256];
  double acc = 0.0;
  for (int g = threadIdx.x; g < groups; g += blockDim.x) {
    const double v = part[(int64_t)g * p + c];
    acc = is_max ? fmax(acc, v) : acc + v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = is_max ? fmax(red[threadIdx.x], red[threadIdx.x + s]) : red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// ------------------------------------------------------------------------------ dense kernels

// Partial Gram C_g = V(rows_g, k0:k0+64)^H diag(w) W(rows_g, :) per workgroup g (row range) and
// column tile k0; a row tile (64 rows of doubles, 32 of complex values) is staged through LDS with
// coalesced loads.
constexpr int kGramCols = 64;
template <class S>
constexpr int gram_rows() {
  return sizeof(S) == sizeof(double) ? 64 : 32;
}

// PC columns of W per pass (4, 8 or 16 by p): wave g of the workgroup takes columns
// g PC/4 .. (g+1) PC/4 - 1 of the pass, lane kk column kk of the V tile.
template <class S, int PC>
__global__ void __launch_bounds__(256)
    k_gram_partial(int64_t n, const S* __restrict__ v, int64_t ldv, int k, const double* __restrict__ wd,
                   const S* __restrict__ w, int64_t ldw, int p, S* __restrict__ part) {
  constexpr int kRows = gram_rows<S>();
  constexpr int kPer = PC / 4;  // W columns per thread
  __shared__ S Vt[kRows][kGramCols + 1];
  __shared__ S Wt[kRows][PC + 1];
  const int k0 = blockIdx.y * kGramCols;
  const int kc = min(kGramCols, k - k0);
  const int tid = threadIdx.x;
  const int kk = tid % 64, cg = tid / 64;  // column of V, group of kPer columns of W
  for (int c0 = 0; c0 < p; c0 += PC) {
    const int pc = min(PC, p - c0);
    S acc[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) acc[q] = zero_of<S>();
    for (int64_t r0 = blockIdx.x * (int64_t)kRows; r0 < n; r0 += (int64_t)gridDim.x * kRows) {
      const int rn = (int)min<int64_t>(kRows, n - r0);
      for (int idx = tid; idx < kRows * kGramCols; idx += 256) {
        const int r = idx % kRows, j = idx / kRows;
        Vt[r][j] = (r < rn && j < kc) ? conj_of(v[r0 + r + (int64_t)(k0 + j) * ldv]) : zero_of<S>();
      }
      for (int idx = tid; idx < kRows * PC; idx += 256) {
        const int r = idx % kRows, c = idx / kRows;
        S val = zero_of<S>();
        if (r < rn && c < pc) {
          val = w[r0 + r + (int64_t)(c0 + c) * ldw];
          if (wd) val = wd[r0 + r] * val;
        }
        Wt[r][c] = val;
      }
      __syncthreads();
#pragma unroll 4
      for (int r = 0; r < kRows; ++r) {
        const S a = Vt[r][kk];
#pragma unroll
        for (int q = 0; q < kPer; ++q) acc[q] += a * Wt[r][cg * kPer + q];
      }
      __syncthreads();
    }
    // partial layout: part[g][c][k] with g = blockIdx.x, c in [0, p), k in [0, k)
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int c = cg * kPer + q;
      if (kk < kc && c < pc) part[((int64_t)blockIdx.x * p + c0 + c) * k + k0 + kk] = acc[q];
    }
  }
}

// C(i, c) = sum over g of part[g][c][i] (doubles: a complex partial is reduced as its two parts, k
// counting doubles). A workgroup takes 64 consecutive outputs o = i + c k (lane = output, so each
// group's 64 partials are one coalesced load) and its 16 waves split the groups, wave w summing
// g = w, w + 16, ... in increasing order; the 16 sums are then added in wave order. Fixed order:
// bitwise reproducible.
constexpr int kGramRedWaves = 16;

__global__ void __launch_bounds__(64 * kGramRedWaves)
    k_gram_reduce(int groups, int k, int p, const double* __restrict__ part, double* __restrict__ c, int64_t ldc) {
  __shared__ double red[kGramRedWaves][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t outs = (int64_t)k * p;
  const int64_t o = (int64_t)blockIdx.x * 64 + lane;
  const int64_t oc = o < outs ? o : outs - 1;  // clamped: loads stay in bounds, the store is masked
  const int64_t stride = (int64_t)p * k;
  double s = 0.0;
  int g = wave;
#pragma unroll 1
  for (; g + 3 * kGramRedWaves < groups; g += 4 * kGramRedWaves) {
    const double a0 = part[(int64_t)g * stride + oc];
    const double a1 = part[(int64_t)(g + kGramRedWaves) * stride + oc];
    const double a2 = part[(int64_t)(g + 2 * kGramRedWaves) * stride + oc];
    const double a3 = part[(int64_t)(g + 3 * kGramRedWaves) * stride + oc];
    s = (((s + a0) + a1) + a2) + a3;
  }
  for (; g < groups; g += kGramRedWaves) s += part[(int64_t)g * stride + oc];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && o < outs) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kGramRedWaves; ++w) t += red[w][lane];
    const int i = (int)(o % k), col = (int)(o / k);
    c[i + col * ldc] = t;
  }
}

// W = beta W + alpha V C; one thread per row, NC columns of W at a time (NC = 4, 8 or 16 by p),
// C staged in LDS row by row (C(j, c0..c0+NC) contiguous, read as vectors), four rows of V in
// flight per thread. Each column sums j = 0..k-1 in order.
template <class S, int NC>
__global__ void __launch_bounds__(256)
    k_block_update(int64_t n, const S* __restrict__ v, int64_t ldv, int k, const S* __restrict__ cmat, int64_t ldc,
                   S* __restrict__ w, int64_t ldw, int p, double alpha, double beta) {
  extern __shared__ __align__(16) unsigned char cs_raw[];
  S* cs = reinterpret_cast<S*>(cs_raw);  // k x NC
  for (int c0 = 0; c0 < p; c0 += NC) {
    const int pc = min(NC, p - c0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < k * NC; idx += blockDim.x) {
      const int j = idx / NC, c = idx % NC;
      cs[idx] = c < pc ? cmat[j + (int64_t)(c0 + c) * ldc] : zero_of<S>();
    }
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      S acc[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] = zero_of<S>();
      int j = 0;
      for (; j + 4 <= k; j += 4) {
        S a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = v[i + (int64_t)(j + u) * ldv];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[c] += a[u] * cs[(j + u) * NC + c];
      }
      for (; j < k; ++j) {
        const S a = v[i + (int64_t)j * ldv];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += a * cs[j * NC + c];
      }
      for (int c = 0; c < pc; ++c) {
        S* ww = w + i + (int64_t)(c0 + c) * ldw;
        *ww = beta == 0.0 ? alpha * acc[c] : alpha * acc[c] + beta * (*ww);
      }
    }
  }
}

// out[c] = x_c^H diag(w) y_c per column (w may be NULL), partials per workgroup then an ordered
// sum (a complex partial is reduced as its two parts).
template <class S>
__global__ void __launch_bounds__(256)
    k_column_dots(int64_t n, const S* __restrict__ x, int64_t ldx, const S* __restrict__ y, int64_t ldy, int p,
                  const double* __restrict__ wd, S* __restrict__ part) {
  __shared__ S red[4][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c0 = 0; c0 < p; c0 += 16) {
    const int pc = min(16, p - c0);
    S acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = zero_of<S>();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      const double w = wd ? wd[i] : 1.0;
      for (int c = 0; c < pc; ++c) acc[c] += (conj_of(x[i + (c0 + c) * ldx]) * w) * y[i + (c0 + c) * ldy];
    }
    for (int c = 0; c < pc; ++c) {
      S v = acc[c];
      for (int off = 32; off > 0; off >>= 1) v += shfl_xor_of(v, off);
      if (lane == 0) red[wave][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < pc) {
      S v = zero_of<S>();
      for (int w = 0; w < 4; ++w) v += red[w][threadIdx.x];
      part[(int64_t)blockIdx.x * p + c0 + threadIdx.x] = v;
    }
    __syncthreads();
  }
}

// Y = a X + b Y for p columns.
__global__ void k_axpby(int64_t n, double a, const double* __restrict__ x, int64_t ldx, double b, double* __restrict__ y,
                        int64_t ldy, int p) {
  const int64_t total = n * p;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx % n, c = idx / n;
    double* yy = y + i + c * ldy;
    const double t = a * x[i + c * ldx];
    *yy = b == 0.0 ? t : t + b * (*yy);
  }
}

// x *= 1 / b with b = sqrt(max(dot[0], 0)) read on the device (0 when b is 0), b stored to
// beta_out: the normalization of a Lanczos vector without reading its norm back.
__global__ void k_scale_by_norm(int64_t len, double* __restrict__ x, const double* __restrict__ dot,
                                double* __restrict__ beta_out) {
  const double b = sqrt(fmax(dot[0], 0.0));
  const double inv = b > 0.0 ? 1.0 / b : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *beta_out = b;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = x[i] * inv;
}

// Per column c: sums of |ax - lambda_c bx|^2, |bx|^2 and |x|^2 (pair_metrics), partials per workgroup.
__global__ void __launch_bounds__(256)
    k_pair_sums(int64_t n, const double* __restrict__ ax, const double* __restrict__ bx, const double* __restrict__ x,
                int64_t ld, int p, const double* __restrict__ lambda, double* __restrict__ part) {
  __shared__ double red[4][3];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = 0; c < p; ++c) {
    const double lam = lambda[c];
    double r2 = 0.0, b2 = 0.0, x2 = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      const double a = ax[i + c * ld], b = bx[i + c * ld], xx = x[i + c * ld];
      const double r = a - lam * b;
      r2 += r * r;
      b2 += b * b;
      x2 += xx * xx;
    }
    for (int off = 32; off > 0; off >>= 1) {
      r2 += __shfl_xor(r2, off);
      b2 += __shfl_xor(b2, off);
      x2 += __shfl_xor(x2, off);
    }
    if (lane == 0) {
      red[wave][0] = r2;
      red[wave][1] = b2;
      red[wave][2] = x2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      double v = 0.0;
      for (int w = 0; w < 4; ++w) v += red[w][threadIdx.x];
      part[(int64_t)blockIdx.x * 3 * p + 3 * c + threadIdx.x] = v;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------ launchers

void* Scratch::get(size_t bytes) {
  if (bytes > size_) {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    size_ = 0;
    check(hipMalloc(&ptr_, bytes), "hipMalloc (kernel scratch)");
    size_ = bytes;
  }
  return ptr_;
}

Scratch::~Scratch() {
  if (ptr_) (void)hipFree(ptr_);
}

namespace {

template <class S>
void csr_spmm_t(int64_t n, const int64_t* rp, const int32_t* ci, const S* v, const S* x, int64_t ldx, S* y,
                int64_t ldy, int64_t p, double alpha, double beta, hipStream_t st) {
  if (n <= 0 || p <= 0) return;
  hipLaunchKernelGGL(k_csr_spmm<S>, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, n, rp, ci, v, x, ldx, y, ldy,
                     (int)p, alpha, beta);
  check(hipGetLastError(), "csr_spmm");
}

template <class S>
void diag_spmm_t(int64_t n, const double* d, const S* x, int64_t ldx, S* y, int64_t ldy, int64_t p, double alpha,
                 double beta, hipStream_t st) {
  if (n <= 0 || p <= 0) return;
  hipLaunchKernelGGL(k_diag_spmm<S>, dim3(grid_for(n * p, 256, 16384)), dim3(256), 0, st, n, d, x, ldx, y, ldy,
                     (int)p, alpha, beta);
  check(hipGetLastError(), "diag_spmm");
}

template <class S>
void shifted_residual_t(Scratch& scratch, int64_t n, const int64_t* arp, const int32_t* aci, const S* av,
                        const int64_t* brp, const int32_t* bci, const S* bv, const double* bd, double sigma, const S* x,
                        int64_t ldx, const S* y, int64_t ldy, S* r, int64_t ldr, int64_t p, double* rnorm2,
                        double* omega, hipStream_t st) {
  if (n <= 0 || p <= 0) return;
  const int groups = grid_for(n, kResThreads, 1024);
  double* part = static_cast<double*>(scratch.get(sizeof(double) * 2 * (size_t)groups * p));
  hipLaunchKernelGGL(k_shifted_residual<S>, dim3(groups), dim3(kResThreads), 0, st, n, arp, aci, av, brp, bci, bv, bd,
                     sigma, x, ldx, y, ldy, r, ldr, (int)p, part, part + (size_t)groups * p);
  if (rnorm2) hipLaunchKernelGGL(k_reduce_partials, dim3(p), dim3(256), 0, st, groups, (int)p, part, rnorm2, 0);
  if (omega)
    hipLaunchKernelGGL(k_reduce_partials, dim3(p), dim3(256), 0, st, groups, (int)p, part + (size_t)groups * p, omega,
                       1);
  check(hipGetLastError(), "shifted_residual");
}

// Row groups of the Gram kernel for an n x k V: enough workgroups to fill the chip, each over many
// row tiles (the partial sums take groups * p * k scalars).
template <class S>
int gram_groups(int64_t n, int64_t k) {
  const int kt = (int)((k + kGramCols - 1) / kGramCols);
  return grid_for(n, gram_rows<S>() * 4, std::max(32, 2048 / kt));
}

template <class S>
void gram_launch(S* part, int groups, int64_t n, const S* v, int64_t ldv, int64_t k, const double* wd, const S* w,
                 int64_t ldw, int64_t p, S* c, int64_t ldc, hipStream_t st) {
  constexpr int parts = sizeof(S) / sizeof(double);  // doubles per scalar
  const int kt = (int)((k + kGramCols - 1) / kGramCols);
  if (p <= 4)
    hipLaunchKernelGGL((k_gram_partial<S, 4>), dim3(groups, kt), dim3(256), 0, st, n, v, ldv, (int)k, wd, w, ldw,
                       (int)p, part);
  else if (p <= 8)
    hipLaunchKernelGGL((k_gram_partial<S, 8>), dim3(groups, kt), dim3(256), 0, st, n, v, ldv, (int)k, wd, w, ldw,
                       (int)p, part);
  else
    hipLaunchKernelGGL((k_gram_partial<S, 16>), dim3(groups, kt), dim3(256), 0, st, n, v, ldv, (int)k, wd, w, ldw,
                       (int)p, part);
  hipLaunchKernelGGL(k_gram_reduce, dim3((int)((parts * k * p + 63) / 64)), dim3(64 * kGramRedWaves), 0, st, groups,
                     (int)(parts * k), (int)p, reinterpret_cast<const double*>(part), reinterpret_cast<double*>(c),
                     parts * ldc);
  check(hipGetLastError(), "gram");
}

template <class S>
void gram_t(Scratch& scratch, int64_t n, const S* v, int64_t ldv, int64_t k, const double* wd, const S* w, int64_t ldw,
            int64_t p, S* c, int64_t ldc, hipStream_t st) {
  if (k <= 0 || p <= 0) return;
  const int groups = gram_groups<S>(n, k);
  S* part = static_cast<S*>(scratch.get(sizeof(S) * (size_t)groups * p * k));
  gram_launch<S>(part, groups, n, v, ldv, k, wd, w, ldw, p, c, ldc, st);
}

template <class S>
void column_dots_t(Scratch& scratch, int64_t n, const S* x, int64_t ldx, const S* y, int64_t ldy, int64_t p,
                   const double* wd, S* out, hipStream_t st) {
  if (p <= 0) return;
  constexpr int parts = sizeof(S) / sizeof(double);
  const int groups = grid_for(n, 256, 512);
  S* part = static_cast<S*>(scratch.get(sizeof(S) * (size_t)groups * p));
  hipLaunchKernelGGL(k_column_dots<S>, dim3(groups), dim3(256), 0, st, n, x, ldx, y, ldy, (int)p, wd, part);
  hipLaunchKernelGGL(k_reduce_partials, dim3(parts * p), dim3(256), 0, st, groups, (int)(parts * p),
                     reinterpret_cast<const double*>(part), reinterpret_cast<double*>(out), 0);
  check(hipGetLastError(), "column_dots");
}

template <class S>
void block_update_t(int64_t n, const S* v, int64_t ldv, int64_t k, const S* c, int64_t ldc, S* w, int64_t ldw,
                    int64_t p, double alpha, double beta, hipStream_t st) {
  if (n <= 0 || p <= 0) return;
  if (k <= 0) {
    if (beta != 1.0) diag_spmm_t<S>(n, nullptr, w, ldw, w, ldw, p, beta, 0.0, st);
    return;
  }
  const int nc = p <= 4 ? 4 : (p <= 8 ? 8 : 16);
  const size_t shm = sizeof(S) * (size_t)k * nc;
  if (shm > 150 * 1024) throw std::runtime_error("block_update: k too large");
  const dim3 grid(grid_for(n, 256, 4096)), block(256);
  if (nc == 4)
    hipLaunchKernelGGL((k_block_update<S, 4>), grid, block, shm, st, n, v, ldv, (int)k, c, ldc, w, ldw, (int)p, alpha,
                       beta);
  else if (nc == 8)
    hipLaunchKernelGGL((k_block_update<S, 8>), grid, block, shm, st, n, v, ldv, (int)k, c, ldc, w, ldw, (int)p, alpha,
                       beta);
  else
    hipLaunchKernelGGL((k_block_update<S, 16>), grid, block, shm, st, n, v, ldv, (int)k, c, ldc, w, ldw, (int)p, alpha,
                       beta);
  check(hipGetLastError(), "block_update");
}

const zval* zc(const double* p) { return reinterpret_cast<const zval*>(p); }
zval* zm(double* p) { return reinterpret_cast<zval*>(p); }

}  // namespace

void csr_spmm(int64_t n, const int64_t* rp, const int32_t* ci, const double* v, const double* x, int64_t ldx,
              double* y, int64_t ldy, int64_t p, double alpha, double beta, hipStream_t st) {
  csr_spmm_t<double>(n, rp, ci, v, x, ldx, y, ldy, p, alpha, beta, st);
}

void diag_spmm(int64_t n, const double* d, const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p,
               double alpha, double beta, hipStream_t st) {
  diag_spmm_t<double>(n, d, x, ldx, y, ldy, p, alpha, beta, st);
}

void shifted_residual(Scratch& scratch, int64_t n, const int64_t* arp, const int32_t* aci, const double* av,
                      const int64_t* brp, const int32_t* bci, const double* bv, const double* bd, double sigma,
                      const double* x, int64_t ldx, const double* y, int64_t ldy, double* r, int64_t ldr, int64_t p,
                      double* rnorm2, double* omega, hipStream_t st) {
  shifted_residual_t<double>(scratch, n, arp, aci, av, brp, bci, bv, bd, sigma, x, ldx, y, ldy, r, ldr, p, rnorm2,
                             omega, st);
}

void gram(Scratch& scratch, int64_t n, const double* v, int64_t ldv, int64_t k, const double* wd, const double* w,
          int64_t ldw, int64_t p, double* c, int64_t ldc, hipStream_t st) {
  gram_t<double>(scratch, n, v, ldv, k, wd, w, ldw, p, c, ldc, st);
}

void column_dots(Scratch& scratch, int64_t n, const double* x, int64_t ldx, const double* y, int64_t ldy, int64_t p,
                 const double* wd, double* out, hipStream_t st) {
  column_dots_t<double>(scratch, n, x, ldx, y, ldy, p, wd, out, st);
}

void block_update(int64_t n, const double* v, int64_t ldv, int64_t k, const double* c, int64_t ldc, double* w,
                  int64_t ldw, int64_t p, double alpha, double beta, hipStream_t st) {
  block_update_t<double>(n, v, ldv, k, c, ldc, w, ldw, p, alpha, beta, st);
}

void zcsr_spmm(int64_t n, const int64_t* rp, const int32_t* ci, const double* v, const double* x, int64_t ldx,
               double* y, int64_t ldy, int64_t p, double alpha, double beta, hipStream_t st) {
  csr_spmm_t<zval>(n, rp, ci, zc(v), zc(x), ldx, zm(y), ldy, p, alpha, beta, st);
}

void zdiag_spmm(int64_t n, const double* d, const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p,
                double alpha, double beta, hipStream_t st) {
  diag_spmm_t<zval>(n, d, zc(x), ldx, zm(y), ldy, p, alpha, beta, st);
}

void zshifted_residual(Scratch& scratch, int64_t n, const int64_t* arp, const int32_t* aci, const double* av,
                       const int64_t* brp, const int32_t* bci, const double* bv, const double* bd, double sigma,
                       const double* x, int64_t ldx, const double* y, int64_t ldy, double* r, int64_t ldr, int64_t p,
                       double* rnorm2, double* omega, hipStream_t st) {
  shifted_residual_t<zval>(scratch, n, arp, aci, zc(av), brp, bci, zc(bv), bd, sigma, zc(x), ldx, zc(y), ldy, zm(r),
                           ldr, p, rnorm2, omega, st);
}

void zgram(Scratch& scratch, int64_t n, const double* v, int64_t ldv, int64_t k, const double* wd, const double* w,
           int64_t ldw, int64_t p, double* c, int64_t ldc, hipStream_t st) {
  gram_t<zval>(scratch, n, zc(v), ldv, k, wd, zc(w), ldw, p, zm(c), ldc, st);
}

void zcolumn_dots(Scratch& scratch, int64_t n, const double* x, int64_t ldx, const double* y, int64_t ldy, int64_t p,
                  const double* wd, double* out, hipStream_t st) {
  column_dots_t<zval>(scratch, n, zc(x), ldx, zc(y), ldy, p, wd, zm(out), st);
}

void zblock_update(int64_t n, const double* v, int64_t ldv, int64_t k, const double* c, int64_t ldc, double* w,
                   int64_t ldw, int64_t p, double alpha, double beta, hipStream_t st) {
  block_update_t<zval>(n, zc(v), ldv, k, zc(c), ldc, zm(w), ldw, p, alpha, beta, st);
}

void pair_sums(Scratch& scratch, int64_t n, const double* ax, const double* bx, const double* x, int64_t ld, int64_t p,
               const double* lambda, double* out, hipStream_t st) {
  if (p <= 0) return;
  const int groups = grid_for(n, 256, 256);
  double* part = static_cast<double*>(scratch.get(sizeof(double) * (size_t)groups * 3 * p));
  hipLaunchKernelGGL(k_pair_sums, dim3(groups), dim3(256), 0, st, n, ax, bx, x, ld, (int)p, lambda, part);
  hipLaunchKernelGGL(k_reduce_partials, dim3(3 * p), dim3(256), 0, st, groups, (int)(3 * p), part, out, 0);
  check(hipGetLastError(), "pair_sums");
}

void axpby(int64_t n, double a, const double* x, int64_t ldx, double b, double* y, int64_t ldy, int64_t p,
           hipStream_t st) {
  if (n <= 0 || p <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n * p, 256, 16384)), dim3(256), 0, st, n, a, x, ldx, b, y, ldy, (int)p);
  check(hipGetLastError(), "axpby");
}

void scale_by_norm(int64_t len, double* x, const double* dot, double* beta_out, hipStream_t st) {
  if (len <= 0) return;
  hipLaunchKernelGGL(k_scale_by_norm, dim3(grid_for(len, 256, 4096)), dim3(256), 0, st, len, x, dot, beta_out);
  check(hipGetLastError(), "scale_by_norm");
}

// ------------------------------------------------------------------------------ no-allocation gram

// Partial sums of gram_noalloc: a fixed device array instead of the per-stream arena, so the call
// allocates nothing. Each device that runs the library holds its own copy until the process ends.
// Host code never passes the array's name to a kernel (in host code it is a stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, once per device, kept in
// g_noalloc_ptr.
namespace {
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

}  // namespace hsg

static __device__ __align__(256) double scratch[hsg::kNoAllocScratchDoubles];

namespace hsg {

void* noalloc_scratch() {
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("gram_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress (gram_noalloc scratch)");
    if (!ptr) throw std::runtime_error("gram_noalloc: no device address for the scratch array");
    g_noalloc_ptr[dev] = ptr;
  }
  return g_noalloc_ptr[dev];
}

size_t gram_noalloc_bytes(int64_t n, int64_t k, int64_t p) {
  if (k <= 0 || p <= 0) return 0;
  return sizeof(double) * (size_t)gram_groups<double>(n, k) * (size_t)p * (size_t)k;
}

void gram_noalloc(int64_t n, const double* v, int64_t ldv, int64_t k, const double* wd, const double* w, int64_t ldw,
                  int64_t p, double* c, int64_t ldc, hipStream_t st) {
  if (k <= 0 || p <= 0) return;
  const size_t need = gram_noalloc_bytes(n, k, p);
  if (need > sizeof(double) * (size_t)kNoAllocScratchDoubles)
    throw std::invalid_argument("gram_noalloc: k * p too large for the fixed scratch (" + std::to_string(need) +
                                " bytes needed, " + std::to_string(sizeof(double) * kNoAllocScratchDoubles) +
                                " available)");
  double* part = static_cast<double*>(noalloc_scratch());
  gram_launch<double>(part, gram_groups<double>(n, k), n, v, ldv, k, wd, w, ldw, p, c, ldc, st);
}

}  // namespace hsg
