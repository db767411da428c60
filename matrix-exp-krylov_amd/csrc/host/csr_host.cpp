// CSRHost loaders and row partitioning.
// reference: main.cu:97-245 (MatrixMarket), main.cu:247-523 (Binary CSR v1,
// docs/binary_csr_v1.md), main.cu:526-744 (partition and ghost plans).
// Error texts are the reference's, so the drop-in CLI prints the same lines.
#include "mek/csr_host.hpp"

#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <numeric>
#include <sstream>

namespace {

[[noreturn]] void fail(const std::string& msg) { throw mek::LoadError(msg); }

template <typename F>
void run_or_exit(F&& f) {
  try {
    f();
  } catch (const mek::LoadError& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    std::exit(EXIT_FAILURE);
  }
}

uint32_t le_u32(const unsigned char* p) {
  return static_cast<uint32_t>(p[0]) | (static_cast<uint32_t>(p[1]) << 8) |
         (static_cast<uint32_t>(p[2]) << 16) | (static_cast<uint32_t>(p[3]) << 24);
}

uint64_t le_u64(const unsigned char* p) {
  return static_cast<uint64_t>(le_u32(p)) | (static_cast<uint64_t>(le_u32(p + 4)) << 32);
}

// Build CSR from per-row lists of (col, value) after mirroring: entries of a
// row are ordered by (col, value) and equal columns are summed in that order,
// which is the order the reference's std::sort over (row, col, value) tuples
// produces (main.cu:197-211, :474-486).
struct Triplets {
  std::vector<int> r, c;
  std::vector<double> v;
  void add(int rr, int cc, double vv) { r.push_back(rr); c.push_back(cc); v.push_back(vv); }
};

void triplets_to_csr(const Triplets& t, int rows, std::vector<int>& row_ptr,
                     std::vector<int>& col_idx, std::vector<double>& values) {
  const size_t ne = t.r.size();
  std::vector<long long> start(static_cast<size_t>(rows) + 1, 0);
  for (size_t e = 0; e < ne; ++e) start[static_cast<size_t>(t.r[e]) + 1]++;
  for (int r = 0; r < rows; ++r) start[r + 1] += start[r];
  std::vector<std::pair<int, double>> ent(ne);
  std::vector<long long> fill(start.begin(), start.end() - 1);
  for (size_t e = 0; e < ne; ++e) ent[fill[t.r[e]]++] = {t.c[e], t.v[e]};
  row_ptr.assign(static_cast<size_t>(rows) + 1, 0);
  col_idx.clear();
  values.clear();
  col_idx.reserve(ne);
  values.reserve(ne);
  for (int r = 0; r < rows; ++r) {
    auto b = ent.begin() + start[r], e = ent.begin() + start[r + 1];
    std::sort(b, e);
    for (auto it = b; it != e; ++it) {
      if (static_cast<long long>(col_idx.size()) > row_ptr[r] && col_idx.back() == it->first)
        values.back() += it->second;
      else {
        col_idx.push_back(it->first);
        values.push_back(it->second);
      }
    }
    row_ptr[r + 1] = static_cast<int>(col_idx.size());
  }
}

std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char ch) { return std::tolower(ch); });
  return s;
}

// Line reader over zlib (reads plain and gzip-compressed text alike).
class LineReader {
 public:
  explicit LineReader(const std::string& path) : f_(gzopen(path.c_str(), "rb")) {
    if (f_) gzbuffer(f_, 1 << 20);
  }
  ~LineReader() { if (f_) gzclose(f_); }
  bool ok() const { return f_ != nullptr; }
  bool getline(std::string& line) {
    line.clear();
    char buf[4096];
    for (;;) {
      if (!gzgets(f_, buf, sizeof(buf))) return !line.empty();
      const size_t len = std::strlen(buf);
      line.append(buf, len);
      if (len > 0 && buf[len - 1] == '\n') {
        line.pop_back();
        return true;
      }
    }
  }

 private:
  gzFile f_;
};

}  // namespace

// --------------------------------------------------------- MatrixMarket

void CSRHost::load_matrix_market(const std::string& filename) {
  LineReader in(filename);
  if (!in.ok()) fail("Could not open file " + filename);
  std::string line, object, format, field, symmetry;
  bool header = false;
  while (in.getline(line)) {
    if (line.empty()) continue;
    std::istringstream ss(line);
    std::string tok;
    ss >> tok;
    if (lower(tok) != "%%matrixmarket") continue;
    ss >> object >> format >> field >> symmetry;
    object = lower(object);
    format = lower(format);
    field = lower(field);
    symmetry = lower(symmetry);
    if (object != "matrix") fail("MatrixMarket object must be 'matrix'");
    if (format != "coordinate") fail("Only 'coordinate' format is supported");
    if (field != "real" && field != "pattern") fail("Only 'real' or 'pattern' fields are supported");
    if (symmetry != "general" && symmetry != "symmetric")
      fail("Only 'general' or 'symmetric' symmetry types are supported");
    header = true;
    break;
  }
  if (!header) fail("Failed to parse MatrixMarket header");
  const bool pattern = (field == "pattern");
  const bool symmetric = (symmetry == "symmetric");

  bool dims = false;
  while (in.getline(line)) {
    if (line.empty() || line[0] == '%') continue;
    std::istringstream ss(line);
    if (!(ss >> rows >> cols >> nnz)) fail("Failed to read matrix dimensions");
    dims = true;
    break;
  }
  if (!dims) fail("Matrix dimensions line not found in file " + filename);
  // the reference sizes row_ptr from these without a check (main.cu:166);
  // negative sizes are rejected here instead of reaching the CSR build
  if (rows < 0 || cols < 0 || nnz < 0)
    fail("Invalid matrix dimensions " + std::to_string(rows) + " " + std::to_string(cols) + " " +
         std::to_string(nnz));

  Triplets t;
  const size_t reserve = static_cast<size_t>(std::max<long long>(nnz, 0)) * (symmetric ? 2 : 1);
  t.r.reserve(reserve);
  t.c.reserve(reserve);
  t.v.reserve(reserve);
  long long read = 0;
  while (read < nnz && in.getline(line)) {
    if (line.empty() || line[0] == '%') continue;
    const char* p = line.c_str();
    char* end = nullptr;
    errno = 0;
    const long r1 = std::strtol(p, &end, 10);
    bool ok = end != p;
    p = end;
    const long c1 = ok ? std::strtol(p, &end, 10) : 0;
    ok = ok && end != p;
    if (!ok) fail("Failed to read row/col for entry " + std::to_string(read));
    p = end;
    double val = 1.0;
    if (!pattern) {
      val = std::strtod(p, &end);
      if (end == p) fail("Failed to read value for entry " + std::to_string(read));
    }
    const long r = r1 - 1, c = c1 - 1;
    if (r < 0 || r >= rows || c < 0 || c >= cols)
      fail("MatrixMarket index out of bounds r=" + std::to_string(r + 1) + " c=" + std::to_string(c + 1));
    t.add(static_cast<int>(r), static_cast<int>(c), val);
    if (symmetric && r != c) t.add(static_cast<int>(c), static_cast<int>(r), val);
    ++read;
  }
  if (read != nnz)
    fail("Expected " + std::to_string(nnz) + " entries but read " + std::to_string(read));
  triplets_to_csr(t, rows, row_ptr, col_idx, values);
  nnz = static_cast<long long>(values.size());
  source_value_dtype = 2;
}

void CSRHost::from_matrix_market(const std::string& filename) {
  run_or_exit([&] { load_matrix_market(filename); });
}

// ------------------------------------------------------ Binary CSR v1

void CSRHost::load_binary_csr(const std::string& filename) {
  std::FILE* f = std::fopen(filename.c_str(), "rb");
  if (!f) fail("Could not open binary CSR file " + filename);
  struct Closer { std::FILE* f; ~Closer() { std::fclose(f); } } closer{f};
  std::fseek(f, 0, SEEK_END);
  const long long file_size = static_cast<long long>(std::ftell(f));
  std::fseek(f, 0, SEEK_SET);
  if (file_size < 64) fail("Binary CSR file too small (expected at least 64 bytes): " + filename);
  unsigned char hdr[64];
  if (std::fread(hdr, 1, 64, f) != 64) fail("Failed to read 64-byte header from " + filename);
  static const unsigned char magic[8] = {'C', 'S', 'R', 0, 0, 0, 0, 1};
  if (std::memcmp(hdr, magic, 8) != 0) fail("Invalid Binary CSR magic in " + filename);
  const uint32_t version = le_u32(hdr + 8);
  const uint32_t flags = le_u32(hdr + 12);
  const uint64_t nr = le_u64(hdr + 16), nc = le_u64(hdr + 24), nz = le_u64(hdr + 32);
  const uint32_t idt = le_u32(hdr + 40), vdt = le_u32(hdr + 44);
  const uint64_t res0 = le_u64(hdr + 48), res1 = le_u64(hdr + 56);
  if (version != 1)
    fail("Unsupported Binary CSR version " + std::to_string(version) + " in " + filename);
  if (idt != 1 && idt != 2) fail("Unsupported index_dtype=" + std::to_string(idt) + " in " + filename);
  if (vdt != 1 && vdt != 2) fail("Unsupported value_dtype=" + std::to_string(vdt) + " in " + filename);
  if (res0 != 0 || res1 != 0) fail("Reserved header fields must be zero in " + filename);
  const uint64_t imax = static_cast<uint64_t>(std::numeric_limits<int>::max());
  if (nr > imax || nc > imax || nz > imax)
    fail("Matrix dimensions/nnz exceed int32 limits used by solver in " + filename);
  const uint64_t isz = (idt == 1) ? 4 : 8, vsz = (vdt == 1) ? 4 : 8;
  const uint64_t expected = 64 + (nr + 1) * isz + nz * isz + nz * vsz;
  if (expected != static_cast<uint64_t>(file_size))
    fail("Binary CSR size mismatch in " + filename + " (expected " + std::to_string(expected) +
         " bytes, got " + std::to_string(file_size) + ")");

  // indices fit in int32 after the checks above (indptr <= nnz, col < ncols)
  auto read_index = [&](std::vector<int64_t>& dst, uint64_t count, const char* what) {
    dst.resize(count);
    if (idt == 1) {
      std::vector<uint32_t> tmp(count);
      if (count && std::fread(tmp.data(), 4, count, f) != count)
        fail(std::string("Failed to read ") + what + "(u32) from " + filename);
      for (uint64_t i = 0; i < count; ++i) dst[i] = tmp[i];
    } else {
      std::vector<uint64_t> tmp(count);
      if (count && std::fread(tmp.data(), 8, count, f) != count)
        fail(std::string("Failed to read ") + what + "(u64) from " + filename);
      for (uint64_t i = 0; i < count; ++i)
        dst[i] = tmp[i] > static_cast<uint64_t>(INT64_MAX) ? INT64_MAX : static_cast<int64_t>(tmp[i]);
    }
  };
  std::vector<int64_t> ip, ix;
  read_index(ip, nr + 1, "indptr");
  read_index(ix, nz, "indices");
  std::vector<double> data(nz);
  if (vdt == 1) {
    std::vector<float> tmp(nz);
    if (nz && std::fread(tmp.data(), 4, nz, f) != nz) fail("Failed to read data(f32) from " + filename);
    for (uint64_t i = 0; i < nz; ++i) data[i] = static_cast<double>(tmp[i]);
  } else {
    if (nz && std::fread(data.data(), 8, nz, f) != nz) fail("Failed to read data(f64) from " + filename);
  }

  rows = static_cast<int>(nr);
  cols = static_cast<int>(nc);
  nnz = static_cast<long long>(nz);
  const int64_t nzs = static_cast<int64_t>(nz);
  if (ip[0] != 0) fail("indptr[0] must be 0 in " + filename);
  for (int r = 0; r < rows; ++r)
    if (ip[r] > ip[r + 1])
      fail("indptr is not non-decreasing at row " + std::to_string(r) + " in " + filename);
  if (ip[rows] != nzs) fail("indptr[nrows] != nnz in " + filename);
  for (int64_t k = 0; k < nzs; ++k)
    if (ix[k] < 0 || static_cast<uint64_t>(ix[k]) >= nc)
      fail("Column index out of range at k=" + std::to_string(k) + " in " + filename);

  const bool sorted = flags & 1u, nodup = flags & 2u, symu = flags & 4u;
  std::ostringstream hs;
  hs << "Binary CSR header: version=" << version << " flags=0x" << std::hex << flags << std::dec
     << " index_dtype=" << (idt == 1 ? "u32" : "u64") << " value_dtype=" << (vdt == 1 ? "f32" : "f64")
     << " symmetric_upper=" << (symu ? "yes" : "no");
  std::cout << hs.str() << std::endl;

  if (sorted || nodup) {
    for (int r = 0; r < rows; ++r) {
      for (int64_t p = ip[r] + 1; p < ip[r + 1]; ++p) {
        if (sorted && ix[p - 1] > ix[p])
          fail("Row " + std::to_string(r) + " is not sorted but sorted flag is set in " + filename);
        if (nodup && ix[p - 1] == ix[p])
          fail("Row " + std::to_string(r) + " has duplicates but no_duplicates flag is set in " + filename);
      }
    }
  }
  source_value_dtype = static_cast<int>(vdt);
  if (symu) {
    if (rows != cols) fail("symmetric_upper requires square matrix in " + filename);
    Triplets t;
    t.r.reserve(2 * nz);
    t.c.reserve(2 * nz);
    t.v.reserve(2 * nz);
    for (int r = 0; r < rows; ++r) {
      for (int64_t p = ip[r]; p < ip[r + 1]; ++p) {
        const int c = static_cast<int>(ix[p]);
        t.add(r, c, data[p]);
        if (c != r) t.add(c, r, data[p]);
      }
    }
    triplets_to_csr(t, rows, row_ptr, col_idx, values);
    nnz = static_cast<long long>(values.size());
  } else {
    row_ptr.resize(static_cast<size_t>(rows) + 1);
    col_idx.resize(nz);
    values = std::move(data);
    for (int r = 0; r <= rows; ++r) row_ptr[r] = static_cast<int>(ip[r]);
    for (int64_t k = 0; k < nzs; ++k) col_idx[k] = static_cast<int>(ix[k]);
  }
}

void CSRHost::from_binary_csr(const std::string& filename) {
  run_or_exit([&] { load_binary_csr(filename); });
}

// ----------------------------------------------------------- partitions

std::vector<CSRHost> CSRHost::partition_rows(int num_gpus) && {
  if (num_gpus != 1) return static_cast<const CSRHost&>(*this).partition_rows(num_gpus);
  std::vector<CSRHost> parts(1);
  parts[0] = std::move(*this);  // row_ptr already starts at 0: nothing to re-base
  row_ptr.clear();
  col_idx.clear();
  values.clear();
  nnz = 0;
  return parts;
}

std::vector<CSRHost> CSRHost::partition_rows(int num_gpus) const& {
  if (num_gpus < 1) throw std::invalid_argument("partition_rows: num_gpus must be >= 1");
  std::vector<CSRHost> parts(num_gpus);
  const int per = rows / num_gpus, rem = rows % num_gpus;
  int start = 0;
  for (int i = 0; i < num_gpus; ++i) {
    const int nr = per + (i < rem ? 1 : 0);
    CSRHost& p = parts[i];
    p.rows = nr;
    p.cols = cols;
    p.source_value_dtype = source_value_dtype;
    const int base = (nr > 0) ? row_ptr[start] : 0;
    const int stop = (nr > 0) ? row_ptr[start + nr] : 0;
    p.nnz = stop - base;
    p.row_ptr.resize(static_cast<size_t>(nr) + 1);
    for (int r = 0; r <= nr; ++r) p.row_ptr[r] = (nr > 0) ? row_ptr[start + r] - base : 0;
    p.col_idx.assign(col_idx.begin() + base, col_idx.begin() + stop);
    p.values.assign(values.begin() + base, values.begin() + stop);
    start += nr;
  }
  return parts;
}

CSRHost::GhostMap CSRHost::build_owner_ghost_maps(const std::vector<CSRHost>& parts, int me,
                                                  int num_gpus) const {
  GhostMap g;
  g.send_offsets.assign(num_gpus, 0);
  g.send_counts.assign(num_gpus, 0);
  g.recv_offsets.assign(num_gpus, 0);
  g.recv_counts.assign(num_gpus, 0);
  std::vector<int> lo(num_gpus), hi(num_gpus);  // [lo, hi) row range per block
  int acc = 0;
  for (int i = 0; i < num_gpus; ++i) {
    lo[i] = acc;
    acc += parts[i].rows;
    hi[i] = acc;
  }
  auto owner_of = [&](int col) {
    const int i = static_cast<int>(std::upper_bound(hi.begin(), hi.end(), col) - hi.begin());
    return i < num_gpus ? i : -1;
  };
  for (int c : parts[me].col_idx) {
    if (c >= lo[me] && c < hi[me]) continue;
    const int o = owner_of(c);
    if (o >= 0) g.owner_to_consumer_map[o].push_back(c);
  }
  size_t off = 0;
  for (int p = 0; p < num_gpus; ++p) {
    if (p == me) continue;
    auto& lst = g.owner_to_consumer_map[p];
    std::sort(lst.begin(), lst.end());
    lst.erase(std::unique(lst.begin(), lst.end()), lst.end());
    if (!lst.empty()) {
      g.incoming_peer_info[p] = {off, lst.size()};
      g.recv_offsets[p] = static_cast<int>(off);
      g.recv_counts[p] = static_cast<int>(lst.size());
      g.incoming_global_col_indices.insert(g.incoming_global_col_indices.end(), lst.begin(), lst.end());
      off += lst.size();
    }
  }
  g.total_recv_size = off;
  for (int p = 0; p < num_gpus; ++p) {
    if (p == me) continue;
    auto& lst = g.consumer_to_owner_map[p];
    for (int c : parts[p].col_idx)
      if (c >= lo[me] && c < hi[me] && !(c >= lo[p] && c < hi[p])) lst.push_back(c);
    std::sort(lst.begin(), lst.end());
    lst.erase(std::unique(lst.begin(), lst.end()), lst.end());
  }
  off = 0;
  for (int p = 0; p < num_gpus; ++p) {
    if (p == me) continue;
    auto& lst = g.consumer_to_owner_map[p];
    if (!lst.empty()) {
      g.outgoing_peer_info[p] = {off, lst.size()};
      g.send_offsets[p] = static_cast<int>(off);
      g.send_counts[p] = static_cast<int>(lst.size());
      g.outgoing_global_col_indices.insert(g.outgoing_global_col_indices.end(), lst.begin(), lst.end());
      off += lst.size();
    }
  }
  g.total_send_size = off;
  return g;
}

void validate_ghost_maps(const std::vector<CSRHost::GhostMap>& maps, int num_gpus) {
  for (int owner = 0; owner < num_gpus; ++owner) {
    for (int consumer = 0; consumer < num_gpus; ++consumer) {
      if (owner == consumer) continue;
      std::vector<int> send, recv;
      auto s = maps[owner].consumer_to_owner_map.find(consumer);
      if (s != maps[owner].consumer_to_owner_map.end()) send = s->second;
      auto r = maps[consumer].owner_to_consumer_map.find(owner);
      if (r != maps[consumer].owner_to_consumer_map.end()) recv = r->second;
      std::sort(send.begin(), send.end());
      send.erase(std::unique(send.begin(), send.end()), send.end());
      std::sort(recv.begin(), recv.end());
      recv.erase(std::unique(recv.begin(), recv.end()), recv.end());
      if (send != recv) {
        std::cerr << "[ERROR] Ghost map mismatch between owner " << owner << " and consumer "
                  << consumer << std::endl;
        std::cerr << "  send size: " << send.size() << " recv size: " << recv.size() << std::endl;
        std::exit(EXIT_FAILURE);
      }
    }
  }
}

double CSRHost::max_abs_value() const {
  double m = 0.0;
  for (double v : values) m = std::max(m, std::abs(v));
  return m;
}

long long CSRHost::f32_inexact(double* max_rel, long long* nonfinite) const {
  long long count = 0, bad = 0;
  double worst = 0.0;
  for (long long k = 0; k < nnz; ++k) {
    const double a = values[k];
    const double r = static_cast<double>(static_cast<float>(a));
    if (!std::isfinite(a) || !std::isfinite(r)) {
      // NaN or infinite values, and finite values beyond FP32's range: counted
      // apart so that max_rel stays a finite relative change
      ++count;
      ++bad;
    } else if (r != a) {
      ++count;
      worst = std::max(worst, std::abs(r - a) / std::abs(a));
    }
  }
  if (max_rel) *max_rel = worst;
  if (nonfinite) *nonfinite = bad;
  return count;
}

double CSRHost::norm_1() const {
  std::vector<double> s(static_cast<size_t>(std::max(cols, 0)), 0.0);
  for (long long k = 0; k < nnz; ++k) s[col_idx[k]] += std::abs(values[k]);
  double m = 0.0;
  for (double v : s) m = std::max(m, v);
  return m;
}

double CSRHost::norm_inf() const {
  double m = 0.0;
  for (int r = 0; r < rows; ++r) {
    double s = 0.0;
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s += std::abs(values[k]);
    m = std::max(m, s);
  }
  return m;
}
