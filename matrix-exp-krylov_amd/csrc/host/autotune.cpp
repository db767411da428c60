// Launch-shape autotuner: times candidate shapes on the device with HIP events.
#include "mek/autotune.hpp"

#include <algorithm>
#include <iterator>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

#include "../kernels/launch.hpp"
#include "hip_check.hpp"
#include "mek/device_context.hpp"

namespace mek {

int time_candidates(const std::vector<int>& candidates, const std::function<void(int)>& launch,
                    hipStream_t s, int reps, std::vector<double>* us_out) {
  if (candidates.empty()) throw Error("ERROR: autotune needs at least one candidate");
  hipEvent_t e0, e1;
  MEK_HIP(hipEventCreate(&e0));
  MEK_HIP(hipEventCreate(&e1));
  std::vector<double> med(candidates.size());
  try {
    for (size_t c = 0; c < candidates.size(); ++c) {
      launch(candidates[c]);  // warm-up
      MEK_HIP_LAUNCHED();
      std::vector<double> t(reps);
      for (int r = 0; r < reps; ++r) {
        MEK_HIP(hipEventRecord(e0, s));
        launch(candidates[c]);
        MEK_HIP(hipEventRecord(e1, s));
        MEK_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        MEK_HIP(hipEventElapsedTime(&ms, e0, e1));
        t[r] = 1000.0 * ms;
      }
      std::sort(t.begin(), t.end());
      med[c] = t[reps / 2];
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (us_out) *us_out = med;
  return static_cast<int>(std::min_element(med.begin(), med.end()) - med.begin());
}

namespace {

using Key = std::tuple<int, int, long long, int, int, int>;  // device, rows, nnz, m, values_f32, layout
std::mutex g_mu;
std::map<Key, TuneChoice> g_cache;

void upload_blocks(DeviceContext& dc, const std::vector<int>& blocks) {
  if (dc.d_row_blocks) MEK_HIP(hipFree(dc.d_row_blocks));
  dc.d_row_blocks = nullptr;
  MEK_HIP(hipMalloc(&dc.d_row_blocks, sizeof(int) * blocks.size()));
  MEK_HIP(hipMemcpy(dc.d_row_blocks, blocks.data(), sizeof(int) * blocks.size(), hipMemcpyHostToDevice));
  dc.n_row_blocks = static_cast<int>(blocks.size()) - 1;
}

}  // namespace

TuneChoice autotune_single(DeviceContext& dc, const CSRHost& A, int m) {
  // sliced-ELL and diagonal copies: the SpMV grid only (no row blocks)
  const int layout = dc.n_diags > 0 ? (dc.d_dia_mask ? 4 : 3) : dc.n_slices > 0 ? 2 : 1;
  const bool copy = layout != 1;
  const Key key{dc.device_id, A.rows, A.nnz, m, dc.values_f32 ? 1 : 0, layout};
  std::ostringstream log;
  TuneChoice best;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) best = it->second;
  }
  const bool hit = best.nnz_cap > 0;
  hipStream_t s = dc.stream_compute;
  MEK_HIP(hipSetDevice(dc.device_id));
  if (!hit && copy) {
    // sliced-ELL or diagonal SpMV: the grid only (the row blocks are not used)
    std::vector<int> grids;
    const long long slices = (static_cast<long long>(dc.local_rows) + kWave - 1) / kWave;
    for (int g : {1024, 2048, kMaxGrid})
      if (g <= std::max<long long>(8, (slices / 4 + 7) / 8 * 8)) grids.push_back(g);
    if (grids.empty()) grids.push_back(dc.spmv_grid);
    ArnoldiBufs b = dc.bufs();
    std::vector<double> us;
    const int i = time_candidates(
        grids,
        [&](int g) {
          b.spmv_grid = g;
          launch_spmv_plain(b, b.V, nullptr, b.w, s);
        },
        s, 5, &us);
    for (size_t c = 0; c < grids.size(); ++c)
      log << (layout == 4 ? "dia_const" : layout == 3 ? "dia" : "sell") << " spmv grid=" << grids[c] << " "
          << us[c] << "us; ";
    best.nnz_cap = kSpmvNnzCap;
    best.spmv_grid = grids[i];
    best.spmv_us = us[i];
  }
  if (!hit) {
    // SpMV: row-block cap x grid (grids are multiples of the 8 XCDs and at
    // most the partial-sum capacity)
    const int caps[] = {512, 1024, kSpmvNnzCap};
    double best_spmv = 1e300;
    for (int cap : (copy ? std::vector<int>{} : std::vector<int>(std::begin(caps), std::end(caps)))) {
      const std::vector<int> blocks = build_row_blocks(A.row_ptr.data(), A.rows, cap, kSpmvRowCap);
      upload_blocks(dc, blocks);
      std::vector<int> grids;
      for (int g : {1024, 2048, kMaxGrid})
        if (g <= std::max(8, (dc.n_row_blocks + 7) / 8 * 8)) grids.push_back(g);
      if (grids.empty()) grids.push_back(8);
      ArnoldiBufs b = dc.bufs();
      std::vector<double> us;
      const int i = time_candidates(
          grids,
          [&](int g) {
            b.spmv_grid = g;
            launch_spmv_plain(b, b.V, nullptr, b.w, s);
          },
          s, 5, &us);
      for (size_t c = 0; c < grids.size(); ++c) log << "spmv cap=" << cap << " grid=" << grids[c] << " " << us[c] << "us; ";
      if (us[i] < best_spmv) {
        best_spmv = us[i];
        best.nnz_cap = cap;
        best.spmv_grid = grids[i];
      }
    }
    if (!copy) best.spmv_us = best_spmv;
    // sweeps: the update sweep at a mid-size basis
    {
      ArnoldiBufs b = dc.bufs();
      const int j = std::max(0, std::min(15, m - 1));
      std::vector<int> grids = {1024, 2048, kMaxGrid};
      std::vector<double> us;
      const int i = time_candidates(
          grids,
          [&](int g) {
            b.grid = g;
            launch_update_dot(b, j, s);
          },
          s, 5, &us);
      for (size_t c = 0; c < grids.size(); ++c) log << "sweep j=" << j << " grid=" << grids[c] << " " << us[c] << "us; ";
      best.sweep_grid = grids[i];
      best.sweep_us = us[i];
    }
    best.log = log.str();
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = best;
  } else {
    best.cached = true;
  }
  // apply
  upload_blocks(dc, build_row_blocks(A.row_ptr.data(), A.rows, best.nnz_cap, kSpmvRowCap));
  dc.spmv_grid = copy ? best.spmv_grid : std::min(best.spmv_grid, std::max(8, (dc.n_row_blocks + 7) / 8 * 8));
  dc.grid = best.sweep_grid;
  // the timing runs wrote into the basis and the status block: start clean
  MEK_HIP(hipMemsetAsync(dc.d_small, 0, sizeof(double) * dc.small_doubles(), s));
  MEK_HIP(hipStreamSynchronize(s));
  return best;
}

}  // namespace mek
