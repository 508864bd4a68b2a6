// Host-side driver of the GPU multifrontal factorization: owns the device copy of the symbolic
// analysis, the level schedule and the numeric buffers of one factorization object (the role of
// one MumpsFactorization instance in the reference, src/factor/mumps_factorization.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../host/analysis.hpp"
#include "solve.hpp"
#include "timing.hpp"

namespace hsg {

struct BatchValues;  // the values of a batched factorization launch (front_common.hpp)

struct EngineOptions {
  int panel_width = 64;       // nb of the large-front path (<= 64)
  int small_front_max = 96;   // fronts of order <= this use the one-workgroup LDS kernel
};

struct InertiaResult {
  int64_t neg = 0, zero = 0, pos = 0, two = 0, exact_zero = 0;
  bool singular = false;
};

// The numeric state of one factorization over an analysis (one Engine may serve several, e.g. the
// spectrum slices solved at the same time on different streams): factor panels, contribution
// blocks, D and pivots, inertia counters, the solve phase's inverse panels and work arrays.
struct Numeric {
  Numeric() = default;
  ~Numeric();
  Numeric(const Numeric&) = delete;
  Numeric& operator=(const Numeric&) = delete;
  void free_all();
  void note_stream(hipStream_t st);  // a stream this factorization's kernels were queued on
  std::vector<hipStream_t> streams_;  // streams whose queued work may use the buffers below
  int precision_ = -1;
  int batch_ = 1;  // factorizations held side by side (batched launches); solves use batch 1 only
  bool valid_ = false;
  void *d_L_ = nullptr, *d_CB_ = nullptr, *d_W_ = nullptr, *d_dval_ = nullptr, *d_doff_ = nullptr;
  int8_t* d_ptype_ = nullptr;
  int32_t *d_piv_ = nullptr, *d_ipiv_ = nullptr, *d_bperm_ = nullptr;
  unsigned long long* d_cnt_ = nullptr;
  void *d_Xp_ = nullptr, *d_Y_ = nullptr, *d_U_ = nullptr;  // solve work arrays (pivot numbering, interleaved)
  int64_t ws_cols_ = 0;
  void *d_G_ = nullptr, *d_Gt_ = nullptr, *d_T_ = nullptr;  // inverse panels (solve phase)
  int32_t *d_pos_ = nullptr, *d_rpos_ = nullptr;  // pivot position of each original row / front row
  PullRow* d_zpull_ = nullptr;  // per pivot position: its row's children update rows (the pull map entry)
  void* d_part_ = nullptr;       // partial sums of split product tiles
  bool g_ready_ = false;
};

class Engine {
 public:
  Engine(Symbolic sym, std::vector<int64_t> asm_dst, std::vector<int64_t> ia, std::vector<int64_t> ib,
         const EngineOptions& opt);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  const Symbolic& symbolic() const { return sym_; }
  int64_t small_fronts() const { return n_small_; }
  double update_flops() const { return update_flops_; }  // one factorization's front-update products
  int64_t workspace_entries() const { return w_total_; }
  int64_t large_fronts() const { return n_large_; }

  // Numeric factorization into nm; returns the inertia counts (synchronizes the stream).
  InertiaResult factorize(Numeric& nm, double alpha, double beta, const double* a_values, const double* b_values,
                          int precision, double null_tol, hipStream_t stream);
  // Z (1..64) factorizations of the same pattern, alpha_z A_z + beta_z B_z, in the same launches
  // (grid z = factorization): inertia counts into out[0..Z); the factors are kept side by side in nm
  // and are not solvable (nm is left without a valid factor). Synchronizes the stream.
  void factorize_batch(Numeric& nm, int Z, const double* alphas, const double* betas, const double* const* a_values,
                       const double* const* b_values, int precision, double null_tol, hipStream_t stream,
                       InertiaResult* out);
  // In-place solve of nrhs FP64 columns with the last factorization held in nm.
  void solve(Numeric& nm, double* x, int64_t nrhs, int64_t ldx, hipStream_t stream);
  void release(Numeric& nm);
  static bool has_factor(const Numeric& nm) { return nm.valid_; }

 private:
  struct Level {
    std::vector<int32_t> h_all, h_small, h_large;
    int32_t* d_all = nullptr;
    int n_all = 0;
    int32_t* d_small = nullptr;
    int n_small = 0;
    int small_maxm = 0;
    int32_t* d_large = nullptr;
    int n_large = 0;
    int large_maxm = 0;
    int large_max_ncol = 0;
    int large_max_cb = 0;
    int max_m = 0, max_ncol = 0;  // over all fronts of the level (solve launches)
    int fwd_tiles = 0, fwd_chunks = 0;   // forward split-GEMV grid
    int fwd_split = 0;                   // columns per chunk of the forward products
    int64_t ft_begin = 0, ft_count = 0;  // forward tasks of the level
    int64_t fh_begin = 0, fh_count = 0;  // its split row tiles (their chunk-0 tasks)
    int64_t bt_begin = 0, bt_count = 0;  // backward tasks of the level
    int64_t bh_begin = 0, bh_count = 0;  // its split column tiles (their chunk-0 tasks)
    int ft_cols = kFwdCols;              // G entries per thread and pass of the forward tasks (2 ... 32)
    bool ft_wave = false;                // forward tasks one per wave (every tile small, one chunk)
    int bt_rows = 32;                    // chunk rows per row group of the backward tasks
    int bt_regs = 32;                    // rows per thread of the backward tasks at most (4 ... 32)
    std::vector<double> upd_flops;       // algorithmic flops of the trailing update after panel p
    double cb_flops = 0.0;               // and of the contribution-block product
  };
  struct TaskRange {
    int64_t begin = 0, count = 0;
  };
  template <class T>
  void factorize_t(Numeric& nm, int Z, const BatchValues& bv, double null_tol, hipStream_t stream, InertiaResult* out);
  template <class T>
  void solve_t(Numeric& nm, double* x, int64_t nrhs, int64_t ldx, hipStream_t stream);
  template <class T>
  void build_inverse_panels(Numeric& nm, hipStream_t stream);  // G, Gt from L (first solve after a factorization)
  void plan_solve();                              // host task lists of the solve phase (symbolic)
  void validate_solve_plan() const;
  void alloc_numeric(Numeric& nm, int precision, int batch);
  // entries of the Bunch-Kaufman block permutation array of one factorization
  int64_t bperm_stride() const { return (int64_t)std::max(1, max_large_per_level_) * opt_.panel_width; }
  void ensure_device();  // uploads the analysis on first use (the analysis itself is host-only)

  Symbolic sym_;
  EngineOptions opt_;
  std::vector<int64_t> h_asm_, h_ia_, h_ib_, h_w_off_;
  // solve phase plan: diagonal 64-block pairs, tile-product stages building G, transpose tiles,
  // and the pull map of the forward products (per front row, the children's U rows mapped onto it)
  std::vector<int32_t> h_pairs_, h_ttiles_;
  std::vector<GemmTask> h_tasks_;
  std::vector<FwdTask> h_ftasks_, h_fheads_;
  FwdTask *d_ftasks_ = nullptr, *d_fheads_ = nullptr;
  std::vector<BwdTask> h_btasks_, h_bheads_;
  BwdTask *d_btasks_ = nullptr, *d_bheads_ = nullptr;
  std::vector<PivSplit> h_psplit_;  // per pivot position: its split tile (backward partials)
  PivSplit* d_psplit_ = nullptr;
  std::vector<TaskRange> gstages_;
  std::vector<int64_t> h_cptr_;
  std::vector<int32_t> h_cidx_;
  std::vector<PullRow> h_cpull_;  // per front row: first two children's rows inline, the rest in cidx
  int64_t t_total_ = 0;
  double g_flops_ = 0.0;  // algorithmic flops of the inverse panels: sum of cb ncol^2 + ncol^3 / 3
  int n_pairs_ = 0, n_ttiles_ = 0;
  int32_t *d_pairs_ = nullptr, *d_ttiles_ = nullptr, *d_cidx_ = nullptr;
  PullRow* d_cpull_ = nullptr;
  GemmTask* d_tasks_ = nullptr;
  bool on_device_ = false;
  int64_t nnz_ = 0;
  int64_t n_small_ = 0, n_large_ = 0;
  int64_t panel_total_ = 0, cb_total_ = 0, w_total_ = 0, rows_total_ = 0;
  int max_large_per_level_ = 0;
  double update_flops_ = 0.0;
  // device symbolic arrays
  int32_t *d_first_ = nullptr, *d_ncol_ = nullptr, *d_m_ = nullptr, *d_child_ptr_ = nullptr,
          *d_child_list_ = nullptr, *d_relmap_ = nullptr, *d_rows_orig_ = nullptr;
  int64_t *d_rowoff_ = nullptr, *d_panel_off_ = nullptr, *d_cb_off_ = nullptr, *d_w_off_ = nullptr;
  int64_t *d_asm_ = nullptr, *d_ia_ = nullptr, *d_ib_ = nullptr;
  std::vector<Level> levels_;
  int64_t part_slots_ = 0;       // slots per group of kMaxRhsChunk right-hand sides
  mutable std::mutex device_mu_;  // ensure_device from factorizations running on several threads
};

// Returns the idle cached device blocks of the current device to the driver; the bytes returned.
size_t release_cached_blocks();

// Throws std::runtime_error with the HIP error string on failure.
void hip_check(hipError_t e, const char* what);

}  // namespace hsg
