// Launch-shape autotuner (additive; the reference takes cuSPARSE's and its
// own fixed launch shapes).  Candidate shapes of a kernel are timed with HIP
// events on the device and data they will run on, and the fastest is kept;
// results are cached per (device, problem shape) for the process.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

struct DeviceContext;
struct CSRHost;

namespace mek {

// Median time in microseconds of `reps` launches of `launch(candidate)` after
// one warm-up launch, for each candidate; returns the index of the fastest.
int time_candidates(const std::vector<int>& candidates, const std::function<void(int)>& launch,
                    hipStream_t s, int reps, std::vector<double>* us_out);

struct TuneChoice {
  int nnz_cap = 0;      // SpMV row-block nonzero cap
  int spmv_grid = 0;    // SpMV workgroups
  int sweep_grid = 0;   // Gram-Schmidt sweep workgroups
  double spmv_us = 0.0;
  double sweep_us = 0.0;
  int sstep_grid = 0;   // s-step block sweep workgroups (when s-step is on)
  int sstep_grid_small = 0;  // ... for a cycle's first block when its window is at most 32 columns
  double sstep_us = 0.0;
  bool cached = false;
  std::string log;      // every candidate and its time
};

// Tunes the single-vector solver's SpMV (row-block cap and grid) and sweep
// grid for this matrix and basis size, applies the choice to dc (row blocks
// re-uploaded, grids set) and returns it.  The device context must be
// allocated; the basis contents are overwritten.
TuneChoice autotune_single(DeviceContext& dc, const CSRHost& A, int m);

}  // namespace mek
