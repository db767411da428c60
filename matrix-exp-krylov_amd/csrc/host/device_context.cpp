// DeviceContext for one MI355X.  reference: main.cu:747-1125.
#include "mek/device_context.hpp"

#include <algorithm>
#include <cstring>
#include <iostream>

#include "../kernels/launch.hpp"
#include "hip_check.hpp"

namespace mek {

std::vector<int> build_row_blocks(const int* row_ptr, int n, int nnz_cap, int row_cap) {
  std::vector<int> starts;
  starts.reserve(static_cast<size_t>(n) / 64 + 2);
  int r = 0;
  while (r < n) {
    starts.push_back(r);
    const int k0 = row_ptr[r];
    if (row_ptr[r + 1] - k0 > nnz_cap) {  // long row: a block of its own
      ++r;
      continue;
    }
    int e = r + 1;
    const int lim = std::min(n, r + row_cap);
    while (e < lim && row_ptr[e + 1] - k0 <= nnz_cap) ++e;
    r = e;
  }
  starts.push_back(n);
  return starts;
}

// Blocks of the streaming kernels: up to 4096 (16 per CU, measured to need
// >= 2048 for full HBM bandwidth), at least 4 rows per thread, a multiple of
// the 8 XCDs.
static int streaming_grid(int n) {
  long long g = (static_cast<long long>(n) + kBlock * 4 - 1) / (kBlock * 4);
  g = std::min<long long>(g, kMaxGrid);
  g = std::max<long long>(g, 8);
  g = (g + 7) / 8 * 8;
  return static_cast<int>(g);
}

}  // namespace mek

using mek::ArnoldiBufs;

void DeviceContext::init_handles() { MEK_HIP(hipSetDevice(device_id)); }

void DeviceContext::create_streams() {
  MEK_HIP(hipSetDevice(device_id));
  if (!stream_compute) MEK_HIP(hipStreamCreate(&stream_compute));
}

size_t DeviceContext::small_doubles() const {
  const size_t M = static_cast<size_t>(allocated_m_max);
  // status block [16 scalars | 8 ints | wH (M+4)] then colscale, h1, h2 (M+2
  // each) and H ((M+2)^2)
  return 16 + 4 + (M + 4) + 3 * (M + 2) + (M + 2) * (M + 2);
}

void DeviceContext::alloc_from(const CSRHost& part, int /*global_cols*/, int m_alloc_max, bool store_f32) {
  MEK_HIP(hipSetDevice(device_id));
  local_rows = part.rows;
  local_nnz = part.nnz;
  if (log) std::cout << "GPU " << device_id << " allocating: local_rows=" << local_rows
            << ", local_nnz=" << local_nnz << std::endl;
  hipStream_t s = stream_compute;
  values_f32 = store_f32;
  const size_t nnz = static_cast<size_t>(std::max<long long>(local_nnz, 0));
  MEK_HIP(hipMalloc(&d_row_ptr, sizeof(int) * (static_cast<size_t>(local_rows) + 1)));
  MEK_HIP(hipMalloc(&d_col_idx, sizeof(int) * std::max<size_t>(nnz, 1)));
  MEK_HIP(hipMemcpyAsync(d_row_ptr, part.row_ptr.data(), sizeof(int) * (local_rows + 1),
                         hipMemcpyHostToDevice, s));
  if (nnz) MEK_HIP(hipMemcpyAsync(d_col_idx, part.col_idx.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s));
  if (store_f32) {
    std::vector<float> f(part.values.begin(), part.values.end());
    MEK_HIP(hipMalloc(&d_values, sizeof(float) * std::max<size_t>(nnz, 1)));
    if (nnz) MEK_HIP(hipMemcpy(d_values, f.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
  } else {
    MEK_HIP(hipMalloc(&d_values, sizeof(double) * std::max<size_t>(nnz, 1)));
    if (nnz) MEK_HIP(hipMemcpyAsync(d_values, part.values.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, s));
  }

  allocated_m_max = std::max(1, m_alloc_max);
  ld = (static_cast<long long>(local_rows) + 63) / 64 * 64;
  const size_t ncols = static_cast<size_t>(allocated_m_max) + 2;
  // V, w and y carry ld rows; rows n..ld-1 stay zero (no kernel writes them):
  // the sweeps read rows in pairs and may touch row n when n is odd
  MEK_HIP(hipMalloc(&d_V_m, sizeof(double) * static_cast<size_t>(ld) * ncols));
  MEK_HIP(hipMemsetAsync(d_V_m, 0, sizeof(double) * static_cast<size_t>(ld) * ncols, s));
  MEK_HIP(hipMalloc(&d_w, sizeof(double) * static_cast<size_t>(std::max<long long>(ld, 64))));
  MEK_HIP(hipMemsetAsync(d_w, 0, sizeof(double) * static_cast<size_t>(std::max<long long>(ld, 64)), s));
  MEK_HIP(hipMalloc(&d_y, sizeof(double) * static_cast<size_t>(std::max<long long>(ld, 64))));
  MEK_HIP(hipMemsetAsync(d_y, 0, sizeof(double) * static_cast<size_t>(std::max<long long>(ld, 64)), s));
  MEK_HIP(hipMalloc(&d_small, sizeof(double) * small_doubles()));
  MEK_HIP(hipMemsetAsync(d_small, 0, sizeof(double) * small_doubles(), s));
  grid = mek::streaming_grid(local_rows);
  // partials: m+4 quantities for either the streaming or the SpMV grid
  MEK_HIP(hipMalloc(&d_partials, sizeof(double) * (ncols + 2) * static_cast<size_t>(mek::kMaxGrid)));
  const int kk = allocated_m_max + 2;
  if (kk > mek::kExpmMaxKLds)
    MEK_HIP(hipMalloc(&d_expm_ws, sizeof(double) * 6 * static_cast<size_t>(kk) * kk));
  MEK_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_status), sizeof(double) * (20 + kk + 2)));
  MEK_HIP(hipStreamSynchronize(s));
}

void DeviceContext::create_descriptors(const CSRHost& part, int /*global_cols*/,
                                       const CSRHost::GhostMap& /*ghost_map*/,
                                       int /*global_row_offset*/, int /*m*/) {
  MEK_HIP(hipSetDevice(device_id));
  std::vector<int> blocks =
      mek::build_row_blocks(part.row_ptr.data(), part.rows, mek::kSpmvNnzCap, mek::kSpmvRowCap);
  n_row_blocks = static_cast<int>(blocks.size()) - 1;
  spmv_grid = std::max(1, std::min(n_row_blocks, mek::kMaxGrid));
  const size_t bytes = sizeof(int) * blocks.size();
  MEK_HIP(hipMalloc(&d_row_blocks, bytes));
  MEK_HIP(hipMemcpy(d_row_blocks, blocks.data(), bytes, hipMemcpyHostToDevice));
  if (log) std::cout << "  GPU " << device_id << " SpMV buffer size: " << bytes << " bytes" << std::endl;
}

mek::ArnoldiBufs DeviceContext::bufs() const {
  ArnoldiBufs b{};
  const size_t M = static_cast<size_t>(allocated_m_max);
  b.n = local_rows;
  b.m = allocated_m_max;
  b.row_ptr = d_row_ptr;
  b.col_idx = d_col_idx;
  b.vals = d_values;
  b.vals_f32 = values_f32 ? 1 : 0;
  b.row_blocks = d_row_blocks;
  b.n_row_blocks = n_row_blocks;
  b.V = d_V_m;
  b.ld = ld;
  b.w = d_w;
  b.y = d_y;
  b.scal = d_small;
  b.iscal = reinterpret_cast<int*>(d_small + 16);
  b.wH = d_small + 20;
  b.colscale = b.wH + (M + 4);
  b.h1 = b.colscale + (M + 2);
  b.h2 = b.h1 + (M + 2);
  b.H = b.h2 + (M + 2);
  b.ldh = static_cast<int>(M + 2);
  b.partials = d_partials;
  b.grid = grid;
  b.spmv_grid = spmv_grid;
  return b;
}

void DeviceContext::alloc_sstep(const std::vector<double>& shifts) {
  MEK_HIP(hipSetDevice(device_id));
  if (shifts.size() > static_cast<size_t>(mek::kSsMaxS)) throw mek::Error("ERROR: more than 16 s-step shifts");
  ss_grid = std::max(8, std::min(grid, 1024));
  if (!d_ss_state) MEK_HIP(hipMalloc(&d_ss_state, sizeof(double) * mek::kSsDoubles));
  if (!d_ss_flags) MEK_HIP(hipMalloc(&d_ss_flags, sizeof(int) * mek::kSsNumFlags));
  // sized for the largest grid the launch-shape tuning may pick
  if (!d_ss_partials)
    MEK_HIP(hipMalloc(&d_ss_partials, sizeof(double) * mek::sstep_partials_doubles(mek::kSsMaxGrid)));
  std::vector<double> st(mek::kSsDoubles, 0.0);
  std::copy(shifts.begin(), shifts.end(), st.begin() + mek::kSsTheta);
  MEK_HIP(hipMemcpy(d_ss_state, st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice));
  MEK_HIP(hipMemset(d_ss_flags, 0, sizeof(int) * mek::kSsNumFlags));
}

mek::SsArgs DeviceContext::ss_args() const {
  mek::SsArgs a{};
  const mek::ArnoldiBufs b = bufs();
  a.V = b.V;
  a.ld = b.ld;
  a.rows = (static_cast<long long>(local_rows) + 15) / 16 * 16;
  a.colscale = b.colscale;
  a.H = b.H;
  a.ldh = b.ldh;
  a.iscal = b.iscal;
  a.spmv_partials = b.partials;
  a.spmv_grid = b.spmv_grid;
  a.st = d_ss_state;
  a.fl = d_ss_flags;
  a.partials = d_ss_partials;
  a.grid = ss_grid;
  return a;
}

void DeviceContext::destroy() {
  if (device_id >= 0) (void)hipSetDevice(device_id);
  auto fr = [](auto*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
  };
  fr(d_row_ptr);
  fr(d_col_idx);
  fr(d_values);
  fr(d_row_blocks);
  fr(d_V_m);
  fr(d_w);
  fr(d_y);
  fr(d_small);
  fr(d_partials);
  fr(d_expm_ws);
  fr(d_ss_state);
  fr(d_ss_flags);
  fr(d_ss_partials);
  if (h_status) (void)hipHostFree(h_status);
  h_status = nullptr;
  if (stream_compute) (void)hipStreamDestroy(stream_compute);
  stream_compute = nullptr;
}
