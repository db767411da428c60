// DeviceContext for one MI355X.  reference: main.cu:747-1125.
#include "mek/device_context.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
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

long long sell_layout(const int* row_ptr, int n, std::vector<long long>& ptr, int* max_width) {
  const long long ns = (static_cast<long long>(std::max(n, 0)) + kWave - 1) / kWave;
  ptr.assign(static_cast<size_t>(ns) + 1, 0);
  int wmax = 0;
  for (long long s = 0; s < ns; ++s) {
    int w = 0;
    const long long r1 = std::min<long long>(n, (s + 1) * kWave);
    for (long long r = s * kWave; r < r1; ++r) w = std::max(w, row_ptr[r + 1] - row_ptr[r]);
    wmax = std::max(wmax, w);
    ptr[s + 1] = ptr[s] + static_cast<long long>(w) * kWave;
  }
  if (max_width) *max_width = wmax;
  return ptr[ns];
}

std::vector<int> dia_offsets(const int* row_ptr, const int* col_idx, const double* values, int n) {
  static_assert(kDiaMaxDiags == kDiaDiagCap, "the diagonal-copy SpMV's capacity");
  std::vector<int> offs;
  for (int r = 0; r < n; ++r) {
    const int e0 = row_ptr[r], e1 = row_ptr[r + 1];
    for (int e = e0; e < e1; ++e) {
      if (e > e0 && col_idx[e] <= col_idx[e - 1]) return {};
      if (std::isnan(values[e])) return {};
      const long long o = static_cast<long long>(col_idx[e]) - r;
      // the k-th entry of a row mostly has the k-th offset of the rows before
      const size_t guess = static_cast<size_t>(e - e0);
      if (guess < offs.size() && offs[guess] == o) continue;
      if (std::find(offs.begin(), offs.end(), o) != offs.end()) continue;
      if (offs.size() == static_cast<size_t>(kDiaMaxDiags)) return {};
      offs.push_back(static_cast<int>(o));
    }
  }
  std::sort(offs.begin(), offs.end());
  return offs;
}

std::vector<double> dia_constants(const int* row_ptr, const int* col_idx, const double* values, int n,
                                  const std::vector<int>& offsets, bool f32) {
  const int nd = static_cast<int>(offsets.size());
  if (nd == 0 || nd > kDiaMaxDiags) return {};
  std::vector<uint64_t> bits(nd, 0);
  std::vector<char> seen(nd, 0);
  for (int r = 0; r < n; ++r) {
    for (int e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
      const long long o = static_cast<long long>(col_idx[e]) - r;
      // rows list the diagonals in ascending order: start at the entry's rank
      int d = std::min(e - row_ptr[r], nd - 1);
      while (d > 0 && offsets[d] > o) --d;
      while (d < nd - 1 && offsets[d] < o) ++d;
      if (offsets[d] != o) return {};
      const double v = f32 ? static_cast<double>(static_cast<float>(values[e])) : values[e];
      uint64_t b;
      std::memcpy(&b, &v, sizeof b);
      if (!seen[d]) {
        seen[d] = 1;
        bits[d] = b;
      } else if (bits[d] != b) {
        return {};
      }
    }
  }
  std::vector<double> out(nd);
  for (int d = 0; d < nd; ++d) std::memcpy(&out[d], &bits[d], sizeof(double));  // unseen: +0.0, never read
  return out;
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

bool DeviceContext::build_sell(const CSRHost& part, int format) {
  MEK_HIP(hipSetDevice(device_id));
  free_sell();
  if (format == 1 || local_rows <= 0) return false;
  std::vector<long long> ptr;
  int wmax = 0;
  const long long ne = mek::sell_layout(part.row_ptr.data(), local_rows, ptr, &wmax);
  if (format == 0 &&
      (wmax > mek::kSellMaxWidth || static_cast<double>(ne) > (1.0 + mek::kSellMaxPad) * static_cast<double>(local_nnz)))
    return false;
  const size_t nes = static_cast<size_t>(std::max<long long>(ne, 1));
  MEK_HIP(hipMalloc(&d_sell_ptr, sizeof(long long) * ptr.size()));
  MEK_HIP(hipMalloc(&d_sell_col, sizeof(int) * nes));
  MEK_HIP(hipMalloc(&d_sell_val, (values_f32 ? sizeof(float) : sizeof(double)) * nes));
  MEK_HIP(hipMemcpy(d_sell_ptr, ptr.data(), sizeof(long long) * ptr.size(), hipMemcpyHostToDevice));
  mek::launch_sell_build(d_row_ptr, d_col_idx, d_values, values_f32 ? 1 : 0, local_rows, d_sell_ptr, d_sell_col,
                         d_sell_val, stream_compute);
  MEK_HIP_LAUNCHED();
  MEK_HIP(hipStreamSynchronize(stream_compute));
  n_slices = static_cast<int>(ptr.size() - 1);
  sell_entries = ne;
  sell_width = wmax;
  // four waves per workgroup, each a slice at a time; a multiple of the 8
  // XCDs when there is work for that many
  int g = std::min(mek::kMaxGrid, (n_slices + 15) / 16);
  spmv_grid = g >= 8 ? (g + 7) / 8 * 8 : std::max(1, g);
  spmv_grid = std::min(spmv_grid, mek::kMaxGrid);
  return true;
}

bool DeviceContext::build_dia(const std::vector<int>& offsets, const std::vector<double>& cvals) {
  MEK_HIP(hipSetDevice(device_id));
  free_dia();
  if (offsets.empty() || local_rows <= 0) return false;
  const int nd = static_cast<int>(offsets.size());
  if (nd > mek::kDiaMaxDiags) throw mek::Error("ERROR: more than 16 diagonals");
  if (!cvals.empty() && cvals.size() != offsets.size())
    throw mek::Error("ERROR: one constant per diagonal");
  const long long ldd = (static_cast<long long>(local_rows) + mek::kWave - 1) / mek::kWave * mek::kWave;
  MEK_HIP(hipMalloc(&d_dia_off, sizeof(int) * nd));
  MEK_HIP(hipMemcpy(d_dia_off, offsets.data(), sizeof(int) * nd, hipMemcpyHostToDevice));
  if (cvals.empty()) {
    const size_t vb = values_f32 ? sizeof(float) : sizeof(double);
    MEK_HIP(hipMalloc(&d_dia_val, vb * static_cast<size_t>(ldd) * nd));
    mek::launch_dia_build(d_row_ptr, d_col_idx, d_values, values_f32 ? 1 : 0, local_rows, d_dia_off, nd, ldd,
                          d_dia_val, stream_compute);
  } else {
    const size_t mb = nd <= 8 ? 1 : 2;
    MEK_HIP(hipMalloc(&d_dia_mask, mb * static_cast<size_t>(ldd)));
    MEK_HIP(hipMalloc(&d_dia_cval, sizeof(double) * nd));
    MEK_HIP(hipMemcpy(d_dia_cval, cvals.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
    mek::launch_dia_mask_build(d_row_ptr, d_col_idx, local_rows, d_dia_off, nd, ldd, d_dia_mask, stream_compute);
  }
  MEK_HIP_LAUNCHED();
  MEK_HIP(hipStreamSynchronize(stream_compute));
  n_diags = nd;
  dia_ld = ldd;
  // four waves per workgroup, each 64 rows (or 128) at a time; a multiple of
  // the 8 XCDs when there is work for that many
  const long long slices = (static_cast<long long>(local_rows) + mek::kWave - 1) / mek::kWave;
  int g = static_cast<int>(std::min<long long>(mek::kMaxGrid, (slices + 15) / 16));
  spmv_grid = g >= 8 ? (g + 7) / 8 * 8 : std::max(1, g);
  spmv_grid = std::min(spmv_grid, mek::kMaxGrid);
  return true;
}

void DeviceContext::free_dia() {
  if (d_dia_val) (void)hipFree(d_dia_val);
  if (d_dia_off) (void)hipFree(d_dia_off);
  if (d_dia_mask) (void)hipFree(d_dia_mask);
  if (d_dia_cval) (void)hipFree(d_dia_cval);
  d_dia_val = nullptr;
  d_dia_off = nullptr;
  d_dia_mask = nullptr;
  d_dia_cval = nullptr;
  n_diags = 0;
  dia_ld = 0;
}

int DeviceContext::build_spmv_layout(const CSRHost& part, int format) {
  if (format < 0 || format > 4) throw mek::Error("ERROR: spmv_format must be 0, 1, 2, 3 or 4");
  free_dia();
  free_sell();
  if (format == 1) return 1;
  if (format == 0 || format == 3 || format == 4) {
    const std::vector<int> offs =
        mek::dia_offsets(part.row_ptr.data(), part.col_idx.data(), part.values.data(), local_rows);
    if (format >= 3 && (offs.empty() || local_rows <= 0))
      throw mek::Error("ERROR: the matrix has no diagonal form (more than 16 diagonals, columns not "
                       "ascending within a row, or NaN values)");
    const bool small_pad = !offs.empty() && static_cast<double>(offs.size()) * local_rows <=
                                                (1.0 + mek::kDiaMaxPad) * static_cast<double>(local_nnz);
    if (format >= 3 || small_pad) {
      std::vector<double> cv;
      if (format != 3)
        cv = mek::dia_constants(part.row_ptr.data(), part.col_idx.data(), part.values.data(), local_rows, offs,
                                values_f32);
      if (format == 4 && cv.empty())
        throw mek::Error("ERROR: the matrix has no constant-diagonal form (a diagonal holds two values)");
      build_dia(offs, cv);
      return cv.empty() ? 3 : 4;
    }
  }
  return build_sell(part, format == 2 ? 2 : 0) ? 2 : 1;
}

void DeviceContext::free_sell() {
  if (d_sell_ptr) (void)hipFree(d_sell_ptr);
  if (d_sell_col) (void)hipFree(d_sell_col);
  if (d_sell_val) (void)hipFree(d_sell_val);
  d_sell_ptr = nullptr;
  d_sell_col = nullptr;
  d_sell_val = nullptr;
  n_slices = 0;
  sell_entries = 0;
  sell_width = 0;
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
  b.sell_ptr = d_sell_ptr;
  b.sell_col = d_sell_col;
  b.sell_val = d_sell_val;
  b.n_slices = n_slices;
  b.dia_val = d_dia_val;
  b.dia_off = d_dia_off;
  b.n_diags = n_diags;
  b.dia_ld = dia_ld;
  b.dia_mask = d_dia_mask;
  b.dia_cval = d_dia_cval;
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
  free_sell();
  free_dia();
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
