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

// ------------------------------------------------------------------------------ sparse kernels

// Y = alpha M X + beta Y, one thread per row (rows of Yee pencils hold ~13 entries).
__global__ void k_csr_spmm(int64_t n, const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                           const double* __restrict__ v, const double* __restrict__ x, int64_t ldx,
                           double* __restrict__ y, int64_t ldy, int p, double alpha, double beta) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = rp[i], e = rp[i + 1];
    for (int c0 = 0; c0 < p; c0 += 8) {
      const int pc = min(8, p - c0);
      double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int64_t q = b; q < e; ++q) {
        const double a = v[q];
        const int64_t j = ci[q];
#pragma unroll
        for (int c = 0; c < 8; ++c)
          if (c < pc) acc[c] += a * x[j + (c0 + c) * ldx];
      }
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (c < pc) {
          double* yy = y + i + (c0 + c) * ldy;
          *yy = beta == 0.0 ? alpha * acc[c] : alpha * acc[c] + beta * (*yy);
        }
    }
  }
}

__global__ void k_diag_spmm(int64_t n, const double* __restrict__ d, const double* __restrict__ x, int64_t ldx,
                            double* __restrict__ y, int64_t ldy, int p, double alpha, double beta) {
  const int64_t total = n * p;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx % n, c = idx / n;
    const double dv = d ? d[i] : 1.0;
    double* yy = y + i + c * ldy;
    const double t = alpha * dv * x[i + c * ldx];
    *yy = beta == 0.0 ? t : t + beta * (*yy);
  }
}

// R = B (X + sigma Y) - A Y with per-workgroup partial sums of |R|^2 and maxima of the
// componentwise quotient |R_i| / (|B| |W| + |A| |Y|)_i, W = X + sigma Y (0/0 = 0, r/0 = inf).
constexpr int kResThreads = 256;

__global__ void __launch_bounds__(kResThreads)
    k_shifted_residual(int64_t n, const int64_t* __restrict__ arp, const int32_t* __restrict__ aci,
                       const double* __restrict__ av, const int64_t* __restrict__ brp,
                       const int32_t* __restrict__ bci, const double* __restrict__ bv,
                       const double* __restrict__ bd, double sigma, const double* __restrict__ x, int64_t ldx,
                       const double* __restrict__ y, int64_t ldy, double* __restrict__ r, int64_t ldr, int p,
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
        double ay = 0.0, aay = 0.0;
        for (int64_t q = arp[i]; q < arp[i + 1]; ++q) {
          const double a = av[q], yy = y[aci[q] + cc * ldy];
          ay += a * yy;
          aay += fabs(a) * fabs(yy);
        }
        double bw = 0.0, bbw = 0.0;
        if (brp) {
          for (int64_t q = brp[i]; q < brp[i + 1]; ++q) {
            const int64_t j = bci[q];
            const double w = x[j + cc * ldx] + sigma * y[j + cc * ldy];
            bw += bv[q] * w;
            bbw += fabs(bv[q]) * fabs(w);
          }
        } else {
          const double dv = bd ? bd[i] : 1.0;
          const double w = x[i + cc * ldx] + sigma * y[i + cc * ldy];
          bw = dv * w;
          bbw = fabs(dv) * fabs(w);
        }
        const double ri = bw - ay;
        r[i + cc * ldr] = ri;
        lsum[c] += ri * ri;
        const double s = bbw + aay;
        const double q = s > 0.0 ? fabs(ri) / s : (ri == 0.0 ? 0.0 : INFINITY);
        lmax[c] = (isfinite(ri) && isfinite(s)) ? fmax(lmax[c], q) : INFINITY;
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

// Partial Gram C_g = V(rows_g, k0:k0+64)^T diag(w) W(rows_g, :) per workgroup g (row range) and
// column tile k0; a row tile of 64 rows is staged through LDS with coalesced loads.
constexpr int kGramRows = 64;
constexpr int kGramCols = 64;
constexpr int kGramP = 16;

__global__ void __launch_bounds__(256)
    k_gram_partial(int64_t n, const double* __restrict__ v, int64_t ldv, int k, const double* __restrict__ wd,
                   const double* __restrict__ w, int64_t ldw, int p, double* __restrict__ part) {
  __shared__ double Vt[kGramRows][kGramCols + 1];
  __shared__ double Wt[kGramRows][kGramP + 1];
  const int k0 = blockIdx.y * kGramCols;
  const int kc = min(kGramCols, k - k0);
  const int tid = threadIdx.x;
  const int kk = tid % 64, cg = tid / 64;  // column of V, group of 4 columns of W
  for (int c0 = 0; c0 < p; c0 += kGramP) {
    const int pc = min(kGramP, p - c0);
    double acc[4] = {0, 0, 0, 0};
    for (int64_t r0 = blockIdx.x * (int64_t)kGramRows; r0 < n; r0 += (int64_t)gridDim.x * kGramRows) {
      const int rn = (int)min<int64_t>(kGramRows, n - r0);
      for (int idx = tid; idx < kGramRows * kGramCols; idx += 256) {
        const int r = idx % kGramRows, j = idx / kGramRows;
        Vt[r][j] = (r < rn && j < kc) ? v[r0 + r + (int64_t)(k0 + j) * ldv] : 0.0;
      }
      for (int idx = tid; idx < kGramRows * kGramP; idx += 256) {
        const int r = idx % kGramRows, c = idx / kGramRows;
        double val = 0.0;
        if (r < rn && c < pc) {
          val = w[r0 + r + (int64_t)(c0 + c) * ldw];
          if (wd) val *= wd[r0 + r];
        }
        Wt[r][c] = val;
      }
      __syncthreads();
#pragma unroll 4
      for (int r = 0; r < kGramRows; ++r) {
        const double a = Vt[r][kk];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] += a * Wt[r][cg * 4 + q];
      }
      __syncthreads();
    }
    // partial layout: part[g][c][k] with g = blockIdx.x, c in [0, p), k in [0, k)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = cg * 4 + q;
      if (kk < kc && c < pc) part[((int64_t)blockIdx.x * p + c0 + c) * k + k0 + kk] = acc[q];
    }
  }
}

// C(i, c) = sum over g of part[g][c][i] in fixed order.
__global__ void k_gram_reduce(int groups, int k, int p, const double* __restrict__ part, double* __restrict__ c,
                              int64_t ldc) {
  const int64_t total = (int64_t)k * p;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx % k), col = (int)(idx / k);
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += part[((int64_t)g * p + col) * k + i];
    c[i + col * ldc] = s;
  }
}

// W = beta W + alpha V C; one thread per row, 16 columns of W at a time, C staged in LDS.
__global__ void __launch_bounds__(256)
    k_block_update(int64_t n, const double* __restrict__ v, int64_t ldv, int k, const double* __restrict__ cmat,
                   int64_t ldc, double* __restrict__ w, int64_t ldw, int p, double alpha, double beta) {
  extern __shared__ double cs[];  // k x 16
  for (int c0 = 0; c0 < p; c0 += 16) {
    const int pc = min(16, p - c0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < k * 16; idx += blockDim.x) {
      const int j = idx % k, c = idx / k;
      cs[idx] = c < pc ? cmat[j + (int64_t)(c0 + c) * ldc] : 0.0;
    }
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      double acc[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) acc[c] = 0.0;
      for (int j = 0; j < k; ++j) {
        const double a = v[i + (int64_t)j * ldv];
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[c] += a * cs[j + c * k];
      }
      for (int c = 0; c < pc; ++c) {
        double* ww = w + i + (int64_t)(c0 + c) * ldw;
        *ww = beta == 0.0 ? alpha * acc[c] : alpha * acc[c] + beta * (*ww);
      }
    }
  }
}

// out[c] = x_c^T diag(w) y_c per column (w may be NULL), partials per workgroup then an ordered sum.
__global__ void __launch_bounds__(256)
    k_column_dots(int64_t n, const double* __restrict__ x, int64_t ldx, const double* __restrict__ y, int64_t ldy,
                  int p, const double* __restrict__ wd, double* __restrict__ part) {
  __shared__ double red[4][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c0 = 0; c0 < p; c0 += 16) {
    const int pc = min(16, p - c0);
    double acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      const double w = wd ? wd[i] : 1.0;
      for (int c = 0; c < pc; ++c) acc[c] += x[i + (c0 + c) * ldx] * w * y[i + (c0 + c) * ldy];
    }
    for (int c = 0; c < pc; ++c) {
      double v = acc[c];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red[wave][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < pc) {
      double v = 0.0;
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

void csr_spmm(int64_t n, const int64_t* rp, const int32_t* ci, const double* v, const double* x, int64_t ldx,
              double* y, int64_t ldy, int64_t p, double alpha, double beta, hipStream_t st) {
  if (n <= 0 || p <= 0) return;
  hipLaunchKernelGGL(k_csr_spmm, dim3(grid_for(n, 256, 16384)), dim3(256), 0, st, n, rp, ci, v, x, ldx, y, ldy,
                     (int)p, alpha, beta);
  check(hipGetLastError(), "csr_spmm");
}

void diag_spmm(int64_t n, const double* d, const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p,
               double alpha, double beta, hipStream_t st) {
  if (n <= 0 || p <= 0) return;
  hipLaunchKernelGGL(k_diag_spmm, dim3(grid_for(n * p, 256, 16384)), dim3(256), 0, st, n, d, x, ldx, y, ldy, (int)p,
                     alpha, beta);
  check(hipGetLastError(), "diag_spmm");
}

void shifted_residual(Scratch& scratch, int64_t n, const int64_t* arp, const int32_t* aci, const double* av,
                      const int64_t* brp, const int32_t* bci, const double* bv, const double* bd, double sigma,
                      const double* x, int64_t ldx, const double* y, int64_t ldy, double* r, int64_t ldr, int64_t p,
                      double* rnorm2, double* omega, hipStream_t st) {
  if (n <= 0 || p <= 0) return;
  const int groups = grid_for(n, kResThreads, 1024);
  double* part = static_cast<double*>(scratch.get(sizeof(double) * 2 * (size_t)groups * p));
  hipLaunchKernelGGL(k_shifted_residual, dim3(groups), dim3(kResThreads), 0, st, n, arp, aci, av, brp, bci, bv, bd,
                     sigma, x, ldx, y, ldy, r, ldr, (int)p, part, part + (size_t)groups * p);
  if (rnorm2) hipLaunchKernelGGL(k_reduce_partials, dim3(p), dim3(256), 0, st, groups, (int)p, part, rnorm2, 0);
  if (omega)
    hipLaunchKernelGGL(k_reduce_partials, dim3(p), dim3(256), 0, st, groups, (int)p, part + (size_t)groups * p, omega,
                       1);
  check(hipGetLastError(), "shifted_residual");
}

void gram(Scratch& scratch, int64_t n, const double* v, int64_t ldv, int64_t k, const double* wd, const double* w,
          int64_t ldw, int64_t p, double* c, int64_t ldc, hipStream_t st) {
  if (k <= 0 || p <= 0) return;
  const int kt = (int)((k + kGramCols - 1) / kGramCols);
  // row groups: enough workgroups to fill the chip, each over many row tiles
  int groups = grid_for(n, kGramRows * 4, std::max(32, 2048 / kt));
  double* part = static_cast<double*>(scratch.get(sizeof(double) * (size_t)groups * p * k));
  hipLaunchKernelGGL(k_gram_partial, dim3(groups, kt), dim3(256), 0, st, n, v, ldv, (int)k, wd, w, ldw, (int)p, part);
  hipLaunchKernelGGL(k_gram_reduce, dim3(grid_for(k * p, 256, 1024)), dim3(256), 0, st, groups, (int)k, (int)p, part,
                     c, ldc);
  check(hipGetLastError(), "gram");
}

void column_dots(Scratch& scratch, int64_t n, const double* x, int64_t ldx, const double* y, int64_t ldy, int64_t p,
                 const double* wd, double* out, hipStream_t st) {
  if (p <= 0) return;
  const int groups = grid_for(n, 256, 512);
  double* part = static_cast<double*>(scratch.get(sizeof(double) * (size_t)groups * p));
  hipLaunchKernelGGL(k_column_dots, dim3(groups), dim3(256), 0, st, n, x, ldx, y, ldy, (int)p, wd, part);
  hipLaunchKernelGGL(k_reduce_partials, dim3(p), dim3(256), 0, st, groups, (int)p, part, out, 0);
  check(hipGetLastError(), "column_dots");
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

void block_update(int64_t n, const double* v, int64_t ldv, int64_t k, const double* c, int64_t ldc, double* w,
                  int64_t ldw, int64_t p, double alpha, double beta, hipStream_t st) {
  if (n <= 0 || p <= 0) return;
  if (k <= 0) {
    if (beta != 1.0) diag_spmm(n, nullptr, w, ldw, w, ldw, p, beta, 0.0, st);
    return;
  }
  const size_t shm = sizeof(double) * (size_t)k * 16;
  if (shm > 150 * 1024) throw std::runtime_error("block_update: k too large");
  hipLaunchKernelGGL(k_block_update, dim3(grid_for(n, 256, 4096)), dim3(256), shm, st, n, v, ldv, (int)k, c, ldc, w, ldw,
                     (int)p, alpha, beta);
  check(hipGetLastError(), "block_update");
}

}  // namespace hsg
