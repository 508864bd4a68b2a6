// Shared device-side definitions of the multifrontal factorization and solve kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hsg {

// Bunch-Kaufman growth constant (1 + sqrt(17)) / 8: the 1x1/2x2 pivot choice of Bunch and
// Kaufman (1977), as in LAPACK's *SYTRF; MUMPS applies the same test inside fronts.
constexpr double kBkAlpha = 0.6403882032022076;

// Supernodal structure, device copies (read-only during factorization and solves).
struct DevSym {
  int32_t n;
  int32_t nsn;
  const int32_t* sn_first;
  const int32_t* sn_ncol;
  const int32_t* sn_m;
  const int64_t* sn_rowoff;
  const int64_t* panel_off;
  const int64_t* cb_off;
  const int64_t* w_off;        // W = L D workspace of large fronts (-1: none)
  const int32_t* child_ptr;
  const int32_t* child_list;
  const int32_t* relmap;       // parallel to the row lists
  const int32_t* rows_orig;    // front rows in the caller's numbering
};

// Pivot type codes per pivot position (pivot order = new numbering after local pivoting).
enum : int8_t { kPiv1 = 0, kPiv2First = 1, kPiv2Second = 2, kPivNull = 3 };

// Counter slots of a factorization (unsigned long long, device).
enum { kCntNeg = 0, kCntZero = 1, kCntPos = 2, kCnt2x2 = 3, kCntKnorm = 4, kCntExactZero = 5, kCntSlots = 8 };

// Values of the factorizations of one batched launch (at most kMaxBatch): for factorization z,
// K_z = alpha[z] A_z + beta[z] B_z with A_z's and B_z's values at a[z], b[z] (b[z] may be null when
// the pattern has no B entries). Passed by value as a kernel argument (2 KiB).
constexpr int kMaxBatch = 64;
struct BatchValues {
  const double* a[kMaxBatch];
  const double* b[kMaxBatch];
  double alpha[kMaxBatch];
  double beta[kMaxBatch];
};

template <class T>
struct DevNum {
  T* L;            // panels: m x ncol per front, column-major; diagonal block holds D
  T* CB;           // contribution blocks: cb x cb per front, column-major (lower triangle)
  T* W;            // L D of large fronts
  T* dval;         // D diagonal, pivot order
  T* doff;         // D(j+1, j) of 2x2 pivots at j (first of the pair)
  int8_t* ptype;   // pivot types
  int32_t* piv;    // pivot position (first + i) -> local front row index
  int32_t* ipiv;   // local front row index (first + r, r < ncol) -> pivot position offset i
  unsigned long long* cnt;
  int32_t* bperm = nullptr;  // per (large front of a level, panel column): the panel's local pivots
  // A batch of factorizations of one analysis (pencils sharing a pattern, or shifts of one
  // pencil) keeps factorization z's arrays at these element offsets from the pointers above; the
  // kernels of a batched launch take z = blockIdx.z.
  int64_t zL = 0, zCB = 0, zW = 0, zn = 0, zbp = 0;
  __device__ __forceinline__ DevNum at(int z) const {
    DevNum o = *this;
    if (z) {
      o.L += z * zL;
      o.CB += z * zCB;
      o.W += z * zW;
      o.dval += z * zn;
      o.doff += z * zn;
      o.ptype += z * zn;
      o.piv += z * zn;
      o.ipiv += z * zn;
      o.cnt += z * kCntSlots;
      o.bperm += z * zbp;
    }
    return o;
  }
};

__device__ __forceinline__ double as_double_bits(unsigned long long b) { return __longlong_as_double((long long)b); }

}  // namespace hsg
