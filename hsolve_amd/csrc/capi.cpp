// extern "C" boundary of libhsg (declared in include/hsg.h). Every entry point catches C++
// exceptions and returns an error code; the message is kept per thread.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <complex>
#include <type_traits>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hsg.h"
#include "host/analysis.hpp"
#include "kernels/engine.hpp"
#include "kernels/krylov.hpp"

struct hsg_factor {
  std::shared_ptr<hsg::Engine> engine;  // the analysis, shared by clones
  hsg::Numeric num;                     // this handle's factorization
  hsg_analysis_info info{};
};

namespace {

thread_local std::string g_last_error;

struct InvalidInput : std::runtime_error {
  using std::runtime_error::runtime_error;
};

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return HSG_OK;
  } catch (const InvalidInput& e) {
    return fail(HSG_INVALID_INPUT, e.what());
  } catch (const std::bad_alloc& e) {
    return fail(HSG_OUT_OF_MEMORY_BUDGET, std::string("allocation failed: ") + e.what());
  } catch (const std::exception& e) {
    const std::string m = e.what();
    if (m.find("hipMalloc") != std::string::npos && m.find("out of memory") != std::string::npos)
      return fail(HSG_OUT_OF_MEMORY_BUDGET, m);
    return fail(HSG_INTERNAL, m);
  } catch (...) {
    return fail(HSG_INTERNAL, "unknown exception");
  }
}

// Scratch arenas for the reduction partials of the Krylov kernels, one per stream: kernels on
// different streams may run at the same time, kernels on one stream are ordered. The map is
// guarded by g_scratch_mu; a call holds its arena's mutex while it launches.
struct StreamScratch {
  std::mutex mu;
  hsg::Scratch arena;
};
std::mutex g_scratch_mu;
std::map<hipStream_t, std::unique_ptr<StreamScratch>> g_scratch;
StreamScratch& scratch_for(hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  std::unique_ptr<StreamScratch>& p = g_scratch[st];
  if (!p) p = std::make_unique<StreamScratch>();
  return *p;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

void require(bool ok, const char* msg) {
  if (!ok) throw InvalidInput(msg);
}

}  // namespace

namespace {

// Host side of hsg_block_orthonormalize: Cholesky factors and triangular inverses of p x p
// Hermitian matrices (p is a Krylov block width, at most a few dozen), real or complex.
template <class S>
double sq_abs(const S& x) {
  return std::norm(std::complex<double>(x));
}
inline double conj_of(double x) { return x; }
inline std::complex<double> conj_of(const std::complex<double>& x) { return std::conj(x); }
inline bool finite_of(double x) { return std::isfinite(x); }
inline bool finite_of(const std::complex<double>& x) { return std::isfinite(x.real()) && std::isfinite(x.imag()); }

// Upper Cholesky factor of Hermitian G (row-major p x p): G = R^H R; false when G is not
// numerically positive definite.
template <class S>
bool cholesky_upper(const std::vector<S>& G, int64_t p, std::vector<S>& R) {
  R.assign((size_t)(p * p), S(0));
  for (int64_t j = 0; j < p; ++j) {
    double s = std::real(std::complex<double>(G[j * p + j]));
    for (int64_t i = 0; i < j; ++i) s -= sq_abs(R[i * p + j]);
    if (!(s > 0.0) || !std::isfinite(s)) return false;
    const double d = std::sqrt(s);
    R[j * p + j] = S(d);
    for (int64_t c = j + 1; c < p; ++c) {
      S t = G[j * p + c];
      for (int64_t i = 0; i < j; ++i) t -= conj_of(R[i * p + j]) * R[i * p + c];
      R[j * p + c] = t / d;
    }
  }
  return true;
}

// Inverse of an upper-triangular R (row-major), by columns of back substitution.
template <class S>
std::vector<S> upper_inverse(const std::vector<S>& R, int64_t p) {
  std::vector<S> X((size_t)(p * p), S(0));
  for (int64_t j = 0; j < p; ++j) {
    X[j * p + j] = S(1) / R[j * p + j];
    for (int64_t i = j - 1; i >= 0; --i) {
      S t = S(0);
      for (int64_t l = i + 1; l <= j; ++l) t += R[i * p + l] * X[l * p + j];
      X[i * p + j] = -t / R[i * p + i];
    }
  }
  return X;
}

// A pinned host buffer per thread for the one device-to-host copy of a block step.
struct PinnedBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  void* get(size_t need) {
    if (need > bytes) {
      if (ptr) (void)hipHostFree(ptr);
      ptr = nullptr;
      bytes = 0;
      if (hipHostMalloc(&ptr, need, hipHostMallocDefault) != hipSuccess)
        throw std::runtime_error("hsg_block_orthonormalize: hipHostMalloc failed");
      bytes = need;
    }
    return ptr;
  }
  ~PinnedBuffer() {
    if (ptr) (void)hipHostFree(ptr);
  }
};
thread_local PinnedBuffer g_pinned;
// R^-1 staged for its host-to-device copy, and the event after that copy: the buffer is rewritten
// only once the previous copy from it has run.
thread_local PinnedBuffer g_pinned_rinv;
struct CopyEvent {
  hipEvent_t ev = nullptr;
  bool pending = false;
  ~CopyEvent() {
    if (ev) (void)hipEventDestroy(ev);
  }
};
thread_local CopyEvent g_rinv_copy;

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class S>
int32_t block_orthonormalize(int64_t n, const double* v, int64_t ldv, int64_t k, double* w, int64_t ldw, int64_t p,
                             const int64_t* b_rowptr, const int32_t* b_colind, const double* b_values,
                             const double* b_diag, double breakdown_tol, double* work, double* norms_in,
                             double* coef_out, double* r_out, hipStream_t st) {
  constexpr bool cplx = std::is_same<S, std::complex<double>>::value;
  constexpr int64_t sw = cplx ? 2 : 1;  // doubles per scalar
  const bool general_b = b_rowptr != nullptr;
  // device work: [norms p][C1 k p][C2 k p][G p p][Rinv p p][tmp n p][B w n p]
  double* buf = work;
  double* d_rinv = buf + sw * (p + 2 * k * p + p * p);
  double* tmp = d_rinv + sw * p * p;
  double* bw = tmp + sw * n * p;
  StreamScratch& sc = scratch_for(st);
  std::lock_guard<std::mutex> lock(sc.mu);
  // B-weighted products of this engine's kernels: diagonal weight, or B W formed first
  auto b_rhs = [&](const double* x) -> const double* {
    if (!general_b) return x;
    if (cplx) hsg::zcsr_spmm(n, b_rowptr, b_colind, b_values, x, ldw, bw, n, p, 1.0, 0.0, st);
    else hsg::csr_spmm(n, b_rowptr, b_colind, b_values, x, ldw, bw, n, p, 1.0, 0.0, st);
    return bw;
  };
  const double* wd = general_b ? nullptr : b_diag;
  const int64_t ldr = general_b ? n : ldw;
  auto gram = [&](const double* vv, int64_t ldvv, int64_t kk, const double* rhs, double* c) {
    if (cplx) hsg::zgram(sc.arena, n, vv, ldvv, kk, wd, rhs, ldr, p, c, kk, st);
    else hsg::gram(sc.arena, n, vv, ldvv, kk, wd, rhs, ldr, p, c, kk, st);
  };
  auto update = [&](const double* vv, int64_t ldvv, int64_t kk, const double* c, double* dst, int64_t ldd,
                    double alpha, double beta) {
    if (cplx) hsg::zblock_update(n, vv, ldvv, kk, c, kk, dst, ldd, p, alpha, beta, st);
    else hsg::block_update(n, vv, ldvv, kk, c, kk, dst, ldd, p, alpha, beta, st);
  };
  // the entry norms, two CGS passes against V, the block's Gram matrix
  {
    const double* rhs = b_rhs(w);
    if (cplx) hsg::zcolumn_dots(sc.arena, n, w, ldw, rhs, ldr, p, wd, buf, st);
    else hsg::column_dots(sc.arena, n, w, ldw, rhs, ldr, p, wd, buf, st);
  }
  for (int i = 0; i < (k ? 2 : 0); ++i) {
    double* c = buf + sw * (p + i * k * p);
    gram(v, ldv, k, b_rhs(w), c);
    update(v, ldv, k, c, w, ldw, -1.0, 1.0);
  }
  double* g_dev = buf + sw * (p + 2 * k * p);
  gram(w, ldw, p, b_rhs(w), g_dev);
  const size_t head = (size_t)sw * (size_t)(p + 2 * k * p + p * p) * sizeof(double);
  double* h = static_cast<double*>(g_pinned.get(head));
  hip_ok(hipMemcpyAsync(h, buf, head, hipMemcpyDeviceToHost, st), "hsg_block_orthonormalize: copy");
  hip_ok(hipStreamSynchronize(st), "hsg_block_orthonormalize: synchronize");
  const S* hs = reinterpret_cast<const S*>(h);
  for (int64_t j = 0; j < p; ++j) norms_in[j] = std::sqrt(std::max(std::real(std::complex<double>(hs[j])), 0.0));
  // coef (k x p, row-major) = C1 + C2, with C stored column-major k x p
  S* coef = reinterpret_cast<S*>(coef_out);
  for (int64_t i = 0; i < k; ++i)
    for (int64_t j = 0; j < p; ++j) {
      S t = S(0);
      for (int pass = 0; pass < 2; ++pass) t += hs[p + pass * k * p + j * k + i];
      coef[i * p + j] = t;
    }
  // G (row-major) = the Gram matrix made exactly Hermitian
  auto gram_host = [&](const S* g) {
    std::vector<S> G((size_t)(p * p));
    for (int64_t i = 0; i < p; ++i)
      for (int64_t j = 0; j < p; ++j) G[i * p + j] = 0.5 * (g[j * p + i] + conj_of(g[i * p + j]));
    return G;
  };
  std::vector<S> G = gram_host(hs + p + 2 * k * p);
  for (int64_t j = 0; j < p; ++j) {
    const double gjj = std::real(std::complex<double>(G[j * p + j]));
    const double floor = breakdown_tol * norms_in[j];
    if (!(gjj > floor * floor)) return 1;
  }
  for (const S& x : G)
    if (!finite_of(x)) return 1;
  std::vector<S> R((size_t)(p * p), S(0));
  for (int64_t j = 0; j < p; ++j) R[j * p + j] = S(1);
  S* hrinv = static_cast<S*>(g_pinned_rinv.get(sizeof(S) * p * p));
  for (int it = 0; it < 2; ++it) {
    std::vector<S> Rk;
    if (!cholesky_upper(G, p, Rk)) return it + 1;
    double dmin = INFINITY, dmax = 0.0;
    for (int64_t j = 0; j < p; ++j) {
      const double d = std::abs(std::complex<double>(Rk[j * p + j]));
      dmin = std::min(dmin, d);
      dmax = std::max(dmax, d);
    }
    if (dmin <= 1e-7 * dmax) return it + 1;  // too ill-conditioned for Cholesky-QR
    const std::vector<S> X = upper_inverse(Rk, p);
    // W <- W X: the update takes C column-major with C[i + j p] = X[i][j]
    if (g_rinv_copy.pending) {
      hip_ok(hipEventSynchronize(g_rinv_copy.ev), "hsg_block_orthonormalize: event");
      g_rinv_copy.pending = false;
    }
    for (int64_t i = 0; i < p; ++i)
      for (int64_t j = 0; j < p; ++j) hrinv[j * p + i] = X[i * p + j];
    hip_ok(hipMemcpyAsync(d_rinv, hrinv, sizeof(S) * p * p, hipMemcpyHostToDevice, st),
           "hsg_block_orthonormalize: copy");
    if (!g_rinv_copy.ev) hip_ok(hipEventCreateWithFlags(&g_rinv_copy.ev, hipEventDisableTiming), "hsg_block_orthonormalize: event");
    hip_ok(hipEventRecord(g_rinv_copy.ev, st), "hsg_block_orthonormalize: event");
    g_rinv_copy.pending = true;
    update(w, ldw, p, d_rinv, tmp, n, 1.0, 0.0);
    hip_ok(hipMemcpy2DAsync(w, sizeof(S) * ldw, tmp, sizeof(S) * n, sizeof(S) * n, p, hipMemcpyDeviceToDevice, st),
           "hsg_block_orthonormalize: copy");
    std::vector<S> RR((size_t)(p * p), S(0));
    for (int64_t i = 0; i < p; ++i)
      for (int64_t j = i; j < p; ++j) {
        S t = S(0);
        for (int64_t l = i; l <= j; ++l) t += Rk[i * p + l] * R[l * p + j];
        RR[i * p + j] = t;
      }
    R.swap(RR);
    if (it == 0) {
      gram(w, ldw, p, b_rhs(w), g_dev);
      hip_ok(hipMemcpyAsync(h, g_dev, sizeof(S) * p * p, hipMemcpyDeviceToHost, st),
             "hsg_block_orthonormalize: copy");
      hip_ok(hipStreamSynchronize(st), "hsg_block_orthonormalize: synchronize");
      G = gram_host(reinterpret_cast<const S*>(h));
    }
  }
  std::memcpy(r_out, R.data(), sizeof(S) * p * p);
  return 0;
}

}  // namespace

extern "C" {

const char* hsg_version(void) { return "0.1.0"; }

const char* hsg_last_error_message(void) { return g_last_error.c_str(); }

int hsg_analyze(int64_t n, const int64_t* rowptr, const int32_t* colind, const int64_t* a_index,
                const int64_t* b_index, const int32_t* order, const hsg_analysis_options* opt, hsg_factor** out) {
  return guarded([&] {
    require(out != nullptr, "hsg_analyze: out is NULL");
    *out = nullptr;
    require(n > 0 && n < (int64_t(1) << 31), "hsg_analyze: n must be in [1, 2^31)");
    require(rowptr && colind && a_index && b_index, "hsg_analyze: NULL pattern array");
    const int64_t nnz = rowptr[n];
    require(rowptr[0] == 0 && nnz >= 0, "hsg_analyze: invalid rowptr");
    for (int64_t i = 0; i < n; ++i) require(rowptr[i + 1] >= rowptr[i], "hsg_analyze: rowptr not monotone");
    for (int64_t e = 0; e < nnz; ++e) require(colind[e] >= 0 && colind[e] < n, "hsg_analyze: column index out of range");
    hsg::OrderingOptions oo;
    hsg::AmalgamationOptions ao;
    hsg::EngineOptions eo;
    if (opt) {
      if (opt->leaf_size > 0) oo.leaf_size = opt->leaf_size;
      if (opt->panel_width > 0) eo.panel_width = opt->panel_width;
      if (opt->small_front_max > 0) eo.small_front_max = opt->small_front_max;
    }
    const auto t0 = std::chrono::steady_clock::now();
    hsg::Symbolic sym = hsg::analyze_pattern(static_cast<int32_t>(n), rowptr, colind, oo, ao, order);
    std::vector<int64_t> dst = hsg::assembly_map(sym, static_cast<int32_t>(n), rowptr, colind);
    const auto t1 = std::chrono::steady_clock::now();
    std::vector<int64_t> ia(a_index, a_index + nnz), ib(b_index, b_index + nnz);
    auto f = std::make_unique<hsg_factor>();
    hsg_analysis_info& in = f->info;
    in.n = n;
    in.nnz = nnz;
    in.supernodes = sym.nsn;
    in.levels = sym.nlevels;
    in.max_front = sym.max_front;
    in.max_pivots = sym.max_ncol;
    in.factor_entries = sym.factor_entries;
    in.flops = sym.flops;
    in.panel_bytes_double = 8.0 * static_cast<double>(sym.panel_off[sym.nsn]);
    in.cb_bytes_double = 8.0 * static_cast<double>(sym.cb_total);
    in.analyze_s = std::chrono::duration<double>(t1 - t0).count();
    in.ordering_s = 0.0;
    f->engine = std::make_shared<hsg::Engine>(std::move(sym), std::move(dst), std::move(ia), std::move(ib), eo);
    in.front_rows = static_cast<int64_t>(f->engine->symbolic().rows.size());
    in.small_fronts = f->engine->small_fronts();
    in.large_fronts = f->engine->large_fronts();
    in.update_flops = f->engine->update_flops();
    in.cb_bytes_unshared = 8.0 * static_cast<double>(f->engine->symbolic().cb_sum);
    in.workspace_bytes_double = 8.0 * static_cast<double>(f->engine->workspace_entries());
    *out = f.release();
  });
}

int hsg_analysis_get_info(const hsg_factor* f, hsg_analysis_info* info) {
  return guarded([&] {
    require(f && info, "hsg_analysis_get_info: NULL argument");
    *info = f->info;
  });
}

int hsg_analysis_get_order(const hsg_factor* f, int32_t* order) {
  return guarded([&] {
    require(f && order, "hsg_analysis_get_order: NULL argument");
    const auto& o = f->engine->symbolic().order;
    std::memcpy(order, o.data(), o.size() * sizeof(int32_t));
  });
}

int hsg_analysis_get_fronts(const hsg_factor* f, int32_t* first, int32_t* ncol, int32_t* m, int32_t* parent,
                            int64_t* rowoff, int32_t* rows) {
  return guarded([&] {
    require(f && first && ncol && m && parent && rowoff && rows, "hsg_analysis_get_fronts: NULL argument");
    const hsg::Symbolic& s = f->engine->symbolic();
    std::memcpy(first, s.sn_first.data(), s.nsn * sizeof(int32_t));
    std::memcpy(ncol, s.sn_ncol.data(), s.nsn * sizeof(int32_t));
    std::memcpy(m, s.sn_m.data(), s.nsn * sizeof(int32_t));
    std::memcpy(parent, s.sn_parent.data(), s.nsn * sizeof(int32_t));
    std::memcpy(rowoff, s.sn_rowoff.data(), (s.nsn + 1) * sizeof(int64_t));
    std::memcpy(rows, s.rows.data(), s.rows.size() * sizeof(int32_t));
  });
}

int hsg_factorize(hsg_factor* f, double alpha, double beta, const double* a_values, const double* b_values,
                  int32_t precision, double null_pivot_tol, void* stream, hsg_inertia* inertia, int32_t* singular) {
  return guarded([&] {
    require(f && inertia, "hsg_factorize: NULL argument");
    require(precision == HSG_DOUBLE || precision == HSG_SINGLE, "hsg_factorize: precision must be 0 or 1");
    require(null_pivot_tol >= 0.0 && null_pivot_tol < 1.0, "hsg_factorize: null_pivot_tol must be in [0, 1)");
    require(a_values != nullptr, "hsg_factorize: a_values is NULL");
    const hsg::InertiaResult r =
        f->engine->factorize(f->num, alpha, beta, a_values, b_values, precision, null_pivot_tol, as_stream(stream));
    inertia->neg = r.neg;
    inertia->zero = r.zero;
    inertia->pos = r.pos;
    inertia->valid = r.singular ? 0 : 1;
    inertia->two_by_two = static_cast<int32_t>(r.two);
    if (singular) *singular = r.singular ? 1 : 0;
  });
}

int hsg_factorize_batch(hsg_factor* f, int32_t count, const double* alphas, const double* betas,
                        const double* const* a_values, const double* const* b_values, int32_t precision,
                        double null_pivot_tol, void* stream, hsg_inertia* inertias) {
  return guarded([&] {
    require(f && alphas && betas && a_values && inertias, "hsg_factorize_batch: NULL argument");
    require(count >= 1 && count <= HSG_MAX_BATCH, "hsg_factorize_batch: count must be in [1, 64]");
    require(precision == HSG_DOUBLE || precision == HSG_SINGLE, "hsg_factorize_batch: precision must be 0 or 1");
    require(null_pivot_tol >= 0.0 && null_pivot_tol < 1.0, "hsg_factorize_batch: null_pivot_tol must be in [0, 1)");
    std::vector<const double*> bv(count, nullptr);
    for (int32_t z = 0; z < count; ++z) {
      require(a_values[z] != nullptr, "hsg_factorize_batch: a_values[z] is NULL");
      if (b_values) bv[z] = b_values[z];
    }
    std::vector<hsg::InertiaResult> r(count);
    f->engine->factorize_batch(f->num, count, alphas, betas, a_values, bv.data(), precision, null_pivot_tol,
                               as_stream(stream), r.data());
    for (int32_t z = 0; z < count; ++z) {
      inertias[z].neg = r[z].neg;
      inertias[z].zero = r[z].zero;
      inertias[z].pos = r[z].pos;
      inertias[z].valid = r[z].singular ? 0 : 1;
      inertias[z].two_by_two = static_cast<int32_t>(r[z].two);
    }
  });
}

int hsg_factor_solve(hsg_factor* f, double* x, int64_t nrhs, int64_t ldx, void* stream) {
  return guarded([&] {
    require(f && x, "hsg_factor_solve: NULL argument");
    require(nrhs >= 0, "hsg_factor_solve: nrhs < 0");
    require(ldx >= f->info.n, "hsg_factor_solve: ldx < n");
    if (!hsg::Engine::has_factor(f->num)) throw std::runtime_error("hsg_factor_solve: no usable factorization");
    f->engine->solve(f->num, x, nrhs, ldx, as_stream(stream));
  });
}

int hsg_release_factors(hsg_factor* f) {
  return guarded([&] {
    require(f != nullptr, "hsg_release_factors: NULL argument");
    f->engine->release(f->num);
  });
}

void hsg_factor_destroy(hsg_factor* f) { delete f; }

int hsg_release_cached_memory(int64_t* bytes) {
  return guarded([&] {
    const size_t b = hsg::release_cached_blocks();
    if (bytes) *bytes = (int64_t)b;
  });
}

int hsg_factor_clone(const hsg_factor* src, hsg_factor** out) {
  return guarded([&] {
    require(src != nullptr && out != nullptr, "hsg_factor_clone: NULL argument");
    *out = nullptr;
    auto f = std::make_unique<hsg_factor>();
    f->engine = src->engine;
    f->info = src->info;
    *out = f.release();
  });
}

int hsg_csr_spmm(int64_t n, const int64_t* rowptr, const int32_t* colind, const double* values, const double* x,
                 int64_t ldx, double* y, int64_t ldy, int64_t p, double alpha, double beta, void* stream) {
  return guarded([&] {
    require(ldx >= n && ldy >= n && p >= 0, "hsg_csr_spmm: invalid sizes");
    hsg::csr_spmm(n, rowptr, colind, values, x, ldx, y, ldy, p, alpha, beta, as_stream(stream));
  });
}

int hsg_diag_spmm(int64_t n, const double* d, const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p,
                  double alpha, double beta, void* stream) {
  return guarded([&] {
    require(ldx >= n && ldy >= n && p >= 0, "hsg_diag_spmm: invalid sizes");
    hsg::diag_spmm(n, d, x, ldx, y, ldy, p, alpha, beta, as_stream(stream));
  });
}

int hsg_shifted_residual(int64_t n, const int64_t* a_rowptr, const int32_t* a_colind, const double* a_values,
                         const int64_t* b_rowptr, const int32_t* b_colind, const double* b_values,
                         const double* b_diag, double sigma, const double* x, int64_t ldx, const double* y,
                         int64_t ldy, double* r, int64_t ldr, int64_t p, double* rnorm2, double* omega,
                         void* stream) {
  return guarded([&] {
    require(ldx >= n && ldy >= n && ldr >= n && p >= 0, "hsg_shifted_residual: invalid sizes");
    StreamScratch& sc = scratch_for(as_stream(stream));
    std::lock_guard<std::mutex> lock(sc.mu);
    hsg::shifted_residual(sc.arena, n, a_rowptr, a_colind, a_values, b_rowptr, b_colind, b_values, b_diag, sigma, x,
                          ldx, y, ldy, r, ldr, p, rnorm2, omega, as_stream(stream));
  });
}

int hsg_gram(int64_t n, const double* v, int64_t ldv, int64_t k, const double* w_diag, const double* w, int64_t ldw,
             int64_t p, double* c, int64_t ldc, void* stream) {
  return guarded([&] {
    require(ldv >= n && ldw >= n && ldc >= k && k >= 0 && p >= 0, "hsg_gram: invalid sizes");
    StreamScratch& sc = scratch_for(as_stream(stream));
    std::lock_guard<std::mutex> lock(sc.mu);
    hsg::gram(sc.arena, n, v, ldv, k, w_diag, w, ldw, p, c, ldc, as_stream(stream));
  });
}

int hsg_gram_noalloc(int64_t n, const double* v, int64_t ldv, int64_t k, const double* w_diag, const double* w,
                     int64_t ldw, int64_t p, double* c, int64_t ldc, void* stream) {
  return guarded([&] {
    require(ldv >= n && ldw >= n && ldc >= k && k >= 0 && p >= 0, "hsg_gram_noalloc: invalid sizes");
    require(hsg::gram_noalloc_bytes(n, k, p) <= sizeof(double) * (size_t)hsg::kNoAllocScratchDoubles,
            "hsg_gram_noalloc: k * p exceeds the fixed scratch (16 MiB of partial sums; use hsg_gram)");
    hsg::gram_noalloc(n, v, ldv, k, w_diag, w, ldw, p, c, ldc, as_stream(stream));
  });
}

int hsg_block_update(int64_t n, const double* v, int64_t ldv, int64_t k, const double* c, int64_t ldc, double* w,
                     int64_t ldw, int64_t p, double alpha, double beta, void* stream) {
  return guarded([&] {
    require(ldv >= n && ldw >= n && ldc >= k && k >= 0 && p >= 0, "hsg_block_update: invalid sizes");
    hsg::block_update(n, v, ldv, k, c, ldc, w, ldw, p, alpha, beta, as_stream(stream));
  });
}

int hsg_column_dots(int64_t n, const double* x, int64_t ldx, const double* y, int64_t ldy, int64_t p,
                    const double* w_diag, double* out, void* stream) {
  return guarded([&] {
    require(ldx >= n && ldy >= n && p >= 0, "hsg_column_dots: invalid sizes");
    StreamScratch& sc = scratch_for(as_stream(stream));
    std::lock_guard<std::mutex> lock(sc.mu);
    hsg::column_dots(sc.arena, n, x, ldx, y, ldy, p, w_diag, out, as_stream(stream));
  });
}

int hsg_lanczos_coefficients(int32_t is_complex, int64_t n, const int64_t* rowptr, const int32_t* colind,
                             const double* values, const double* diag, double* basis, int64_t steps, double* coef,
                             double* work, void* stream) {
  return guarded([&] {
    require(n > 0 && steps >= 1, "hsg_lanczos_coefficients: n and steps must be positive");
    require((rowptr && colind && values) || diag, "hsg_lanczos_coefficients: M needs CSR arrays or a diagonal");
    require(basis && coef && work, "hsg_lanczos_coefficients: NULL argument");
    const hipStream_t st = as_stream(stream);
    const bool cx = is_complex != 0;
    const int64_t len = cx ? 2 * n : n;  // doubles per vector
    StreamScratch& sc = scratch_for(st);
    std::lock_guard<std::mutex> lock(sc.mu);
    for (int64_t j = 0; j < steps; ++j) {
      const double* v = basis + j * len;
      double* w = basis + (j + 1) * len;  // the next vector is formed in place
      if (rowptr) {
        if (cx) hsg::zcsr_spmm(n, rowptr, colind, values, v, n, w, n, 1, 1.0, 0.0, st);
        else hsg::csr_spmm(n, rowptr, colind, values, v, n, w, n, 1, 1.0, 0.0, st);
      } else {
        if (cx) hsg::zdiag_spmm(n, diag, v, n, w, n, 1, 1.0, 0.0, st);
        else hsg::diag_spmm(n, diag, v, n, w, n, 1, 1.0, 0.0, st);
      }
      // alpha_j = Re v_j^H w (the real part is the first double of a complex dot)
      if (cx) hsg::zcolumn_dots(sc.arena, n, v, n, w, n, 1, nullptr, work, st);
      else hsg::column_dots(sc.arena, n, v, n, w, n, 1, nullptr, work, st);
      hsg::hip_check(hipMemcpyAsync(coef + j, work, sizeof(double), hipMemcpyDeviceToDevice, st),
                "hsg_lanczos_coefficients: hipMemcpyAsync");
      // full reorthogonalization against v_0..v_j, classical Gram-Schmidt twice
      for (int pass = 0; pass < 2; ++pass) {
        if (cx) {
          hsg::zgram(sc.arena, n, basis, n, j + 1, nullptr, w, n, 1, work, j + 1, st);
          hsg::zblock_update(n, basis, n, j + 1, work, j + 1, w, n, 1, -1.0, 1.0, st);
        } else {
          hsg::gram(sc.arena, n, basis, n, j + 1, nullptr, w, n, 1, work, j + 1, st);
          hsg::block_update(n, basis, n, j + 1, work, j + 1, w, n, 1, -1.0, 1.0, st);
        }
      }
      if (j + 1 == steps) break;
      if (cx) hsg::zcolumn_dots(sc.arena, n, w, n, w, n, 1, nullptr, work, st);
      else hsg::column_dots(sc.arena, n, w, n, w, n, 1, nullptr, work, st);
      hsg::scale_by_norm(len, w, work, coef + steps + j, st);
    }
  });
}

int hsg_zcsr_spmm(int64_t n, const int64_t* rowptr, const int32_t* colind, const double* values, const double* x,
                  int64_t ldx, double* y, int64_t ldy, int64_t p, double alpha, double beta, void* stream) {
  return guarded([&] {
    require(ldx >= n && ldy >= n && p >= 0, "hsg_zcsr_spmm: invalid sizes");
    hsg::zcsr_spmm(n, rowptr, colind, values, x, ldx, y, ldy, p, alpha, beta, as_stream(stream));
  });
}

int hsg_zdiag_spmm(int64_t n, const double* d, const double* x, int64_t ldx, double* y, int64_t ldy, int64_t p,
                   double alpha, double beta, void* stream) {
  return guarded([&] {
    require(ldx >= n && ldy >= n && p >= 0, "hsg_zdiag_spmm: invalid sizes");
    hsg::zdiag_spmm(n, d, x, ldx, y, ldy, p, alpha, beta, as_stream(stream));
  });
}

int hsg_zshifted_residual(int64_t n, const int64_t* a_rowptr, const int32_t* a_colind, const double* a_values,
                          const int64_t* b_rowptr, const int32_t* b_colind, const double* b_values,
                          const double* b_diag, double sigma, const double* x, int64_t ldx, const double* y,
                          int64_t ldy, double* r, int64_t ldr, int64_t p, double* rnorm2, double* omega,
                          void* stream) {
  return guarded([&] {
    require(ldx >= n && ldy >= n && ldr >= n && p >= 0, "hsg_zshifted_residual: invalid sizes");
    StreamScratch& sc = scratch_for(as_stream(stream));
    std::lock_guard<std::mutex> lock(sc.mu);
    hsg::zshifted_residual(sc.arena, n, a_rowptr, a_colind, a_values, b_rowptr, b_colind, b_values, b_diag, sigma,
                           x, ldx, y, ldy, r, ldr, p, rnorm2, omega, as_stream(stream));
  });
}

int hsg_zgram(int64_t n, const double* v, int64_t ldv, int64_t k, const double* w_diag, const double* w, int64_t ldw,
              int64_t p, double* c, int64_t ldc, void* stream) {
  return guarded([&] {
    require(ldv >= n && ldw >= n && ldc >= k && k >= 0 && p >= 0, "hsg_zgram: invalid sizes");
    StreamScratch& sc = scratch_for(as_stream(stream));
    std::lock_guard<std::mutex> lock(sc.mu);
    hsg::zgram(sc.arena, n, v, ldv, k, w_diag, w, ldw, p, c, ldc, as_stream(stream));
  });
}

int hsg_zblock_update(int64_t n, const double* v, int64_t ldv, int64_t k, const double* c, int64_t ldc, double* w,
                      int64_t ldw, int64_t p, double alpha, double beta, void* stream) {
  return guarded([&] {
    require(ldv >= n && ldw >= n && ldc >= k && k >= 0 && p >= 0, "hsg_zblock_update: invalid sizes");
    hsg::zblock_update(n, v, ldv, k, c, ldc, w, ldw, p, alpha, beta, as_stream(stream));
  });
}

int hsg_zcolumn_dots(int64_t n, const double* x, int64_t ldx, const double* y, int64_t ldy, int64_t p,
                     const double* w_diag, double* out, void* stream) {
  return guarded([&] {
    require(ldx >= n && ldy >= n && p >= 0, "hsg_zcolumn_dots: invalid sizes");
    StreamScratch& sc = scratch_for(as_stream(stream));
    std::lock_guard<std::mutex> lock(sc.mu);
    hsg::zcolumn_dots(sc.arena, n, x, ldx, y, ldy, p, w_diag, out, as_stream(stream));
  });
}

int hsg_block_orthonormalize(int32_t is_complex, int64_t n, const double* v, int64_t ldv, int64_t k, double* w,
                             int64_t ldw, int64_t p, const int64_t* b_rowptr, const int32_t* b_colind,
                             const double* b_values, const double* b_diag, double breakdown_tol, double* work,
                             double* norms_in, double* coef, double* r, int32_t* status, void* stream) {
  return guarded([&] {
    require(n >= 0 && k >= 0 && p > 0 && ldw >= n && (k == 0 || ldv >= n) && work && norms_in && r && status &&
                (k == 0 || coef),
            "hsg_block_orthonormalize: invalid arguments");
    require(!b_rowptr || (b_colind && b_values), "hsg_block_orthonormalize: incomplete CSR B");
    *status = is_complex ? block_orthonormalize<std::complex<double>>(n, v, ldv, k, w, ldw, p, b_rowptr, b_colind,
                                                                      b_values, b_diag, breakdown_tol, work, norms_in,
                                                                      coef, r, as_stream(stream))
                         : block_orthonormalize<double>(n, v, ldv, k, w, ldw, p, b_rowptr, b_colind, b_values, b_diag,
                                                        breakdown_tol, work, norms_in, coef, r, as_stream(stream));
  });
}

int hsg_pair_sums(int64_t n, const double* ax, const double* bx, const double* x, int64_t ld, int64_t p,
                  const double* lambda, double* out, void* stream) {
  return guarded([&] {
    require(ld >= n && p >= 0, "hsg_pair_sums: invalid sizes");
    StreamScratch& sc = scratch_for(as_stream(stream));
    std::lock_guard<std::mutex> lock(sc.mu);
    hsg::pair_sums(sc.arena, n, ax, bx, x, ld, p, lambda, out, as_stream(stream));
  });
}

int hsg_axpby(int64_t n, double a, const double* x, int64_t ldx, double b, double* y, int64_t ldy, int64_t p,
              void* stream) {
  return guarded([&] {
    require(ldx >= n && ldy >= n && p >= 0, "hsg_axpby: invalid sizes");
    hsg::axpby(n, a, x, ldx, b, y, ldy, p, as_stream(stream));
  });
}

int hsg_timing_enable(int on) {
  return guarded([&] {
    if (on < 0 || on > 2) throw InvalidInput("timing level must be 0, 1 or 2");
    hsg::timing_enable(on);
  });
}

int hsg_timing_read(int phase, double* ms, int64_t* launches, double* work, int64_t* calls) {
  return guarded([&] {
    if (phase < 0 || phase >= hsg::kPhaseCount) throw InvalidInput("timing phase out of range");
    const hsg::PhaseTotals t = hsg::timing_read(phase);
    if (ms) *ms = t.ms;
    if (launches) *launches = t.launches;
    if (work) *work = t.work;
    if (calls) *calls = t.calls;
  });
}

}  // extern "C"
