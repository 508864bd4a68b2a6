// Launchers of the operator / Krylov kernels (krylov.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace hsg {

// Device scratch that grows on demand (partial sums of the two-pass reductions).
class Scratch {
 public:
  Scratch() = default;
  ~Scratch();
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  void* get(size_t bytes);

 private:
  void* ptr_ = nullptr;
  size_t size_ = 0;
};

void csr_spmm(int64_t n, const int64_t* rp, const int32_t* ci, const double* v, const double* x, int64_t ldx,
              double* y, int64_t ldy, int64_t p, double alpha, double beta, hipStream_t st);
void diag_spmm(int64_t n, const double* d, const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p,
               double alpha, double beta, hipStream_t st);
void shifted_residual(Scratch& scratch, int64_t n, const int64_t* arp, const int32_t* aci, const double* av,
                      const int64_t* brp, const int32_t* bci, const double* bv, const double* bd, double sigma,
                      const double* x, int64_t ldx, const double* y, int64_t ldy, double* r, int64_t ldr, int64_t p,
                      double* rnorm2, double* omega, hipStream_t st);
void gram(Scratch& scratch, int64_t n, const double* v, int64_t ldv, int64_t k, const double* wd, const double* w,
          int64_t ldw, int64_t p, double* c, int64_t ldc, hipStream_t st);
void column_dots(Scratch& scratch, int64_t n, const double* x, int64_t ldx, const double* y, int64_t ldy, int64_t p,
                 const double* wd, double* out, hipStream_t st);
void pair_sums(Scratch& scratch, int64_t n, const double* ax, const double* bx, const double* x, int64_t ld, int64_t p,
               const double* lambda, double* out, hipStream_t st);
// Gram without allocation: the partial sums go to a fixed device array of kNoAllocScratchDoubles
// doubles (16 MiB) per device instead of the per-stream arena; throws std::invalid_argument when
// gram_noalloc_bytes(n, k, p) exceeds it. Calls that overlap on one device share the array.
constexpr int64_t kNoAllocScratchDoubles = int64_t(1) << 21;
constexpr int kMaxDevices = 64;
size_t gram_noalloc_bytes(int64_t n, int64_t k, int64_t p);
void gram_noalloc(int64_t n, const double* v, int64_t ldv, int64_t k, const double* wd, const double* w, int64_t ldw,
                  int64_t p, double* c, int64_t ldc, hipStream_t st);
// x *= 1 / sqrt(max(dot[0], 0)) (0 when that is 0), the square root stored to beta_out (device)
void scale_by_norm(int64_t len, double* x, const double* dot, double* beta_out, hipStream_t st);
void axpby(int64_t n, double a, const double* x, int64_t ldx, double b, double* y, int64_t ldy, int64_t p,
           hipStream_t st);
void block_update(int64_t n, const double* v, int64_t ldv, int64_t k, const double* c, int64_t ldc, double* w,
                  int64_t ldw, int64_t p, double alpha, double beta, hipStream_t st);

// Complex variants: every value, vector and C array holds interleaved (real, imaginary) doubles; the
// diagonal weights d / wd stay real. gram and column_dots conjugate their first operand.
void zcsr_spmm(int64_t n, const int64_t* rp, const int32_t* ci, const double* v, const double* x, int64_t ldx,
               double* y, int64_t ldy, int64_t p, double alpha, double beta, hipStream_t st);
void zdiag_spmm(int64_t n, const double* d, const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p,
                double alpha, double beta, hipStream_t st);
void zshifted_residual(Scratch& scratch, int64_t n, const int64_t* arp, const int32_t* aci, const double* av,
                       const int64_t* brp, const int32_t* bci, const double* bv, const double* bd, double sigma,
                       const double* x, int64_t ldx, const double* y, int64_t ldy, double* r, int64_t ldr, int64_t p,
                       double* rnorm2, double* omega, hipStream_t st);
void zgram(Scratch& scratch, int64_t n, const double* v, int64_t ldv, int64_t k, const double* wd, const double* w,
           int64_t ldw, int64_t p, double* c, int64_t ldc, hipStream_t st);
void zcolumn_dots(Scratch& scratch, int64_t n, const double* x, int64_t ldx, const double* y, int64_t ldy, int64_t p,
                  const double* wd, double* out, hipStream_t st);
void zblock_update(int64_t n, const double* v, int64_t ldv, int64_t k, const double* c, int64_t ldc, double* w,
                   int64_t ldw, int64_t p, double alpha, double beta, hipStream_t st);

}  // namespace hsg
