// Optional phase timing of the engine with HIP events recorded on the stream of each call: per
// phase the summed event time, the kernel launches inside it and the algorithmic work (flops or
// bytes, as defined where the phase is recorded). Off by default; the events are only read
// (synchronized) by timing_read.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hsg {

enum TimedPhase : int {
  kPhaseFactor = 0,    // numeric factorization (assembly, fronts, panels, updates): flops
  kPhaseInverse = 1,   // inverse panels G, Gt from L (first solve after a factorization): flops
  kPhaseForward = 2,   // forward sweep of a solve (products with their gathers, reductions): bytes
  kPhaseBackward = 3,  // backward sweep of a solve (transposed products): bytes
  kPhaseUpdate = 4,    // the front-update products (k_gemm_update launches) of factorizations: flops
                       // of their lower-triangle outputs; recorded per launch at detail level 2 only
  kPhaseCount = 5
};

struct PhaseTotals {
  double ms = 0.0;
  int64_t launches = 0;
  double work = 0.0;
  int64_t calls = 0;
};

// level 0: off, 1: phases, 2: phases and every front-update launch; clears the records either way
void timing_enable(int level);
int timing_level();
// Records the start event of a phase on `st`; returns a token for timing_end (-1: timing off).
int64_t timing_begin(int phase, hipStream_t st);
void timing_end(int64_t token, hipStream_t st, int64_t launches, double work);
PhaseTotals timing_read(int phase);  // synchronizes the recorded events

}  // namespace hsg
