// matrix_exp: drop-in command line of the reference (main.cu:2083-2324).
//
//   matrix_exp [SIZE|PATH] [m] [max_restarts] [--m N] [--t T] [--max-restarts R]
//
// Same argument classification, built-in diagonal test matrix, t clamp to
// 10/max|a_ij|, m default and clamp, ArnoldiParams(m, t, 1e-6, R, true),
// v = 0.1, OUTPUT_Y_FILE and stdout lines.  Additive flags (none changes a
// reference invocation):
//   --tol X              tolerance (reference CLI: fixed 1e-6)
//   --no-t-clamp         use t as given
//   --v-file PATH        start vector, one value per line (default 0.1)
//   --step-control reference|expokit|expokit-rel
//   --gs always|never|dgks
//   --f32-values         store matrix values in FP32 on the device when every
//                        value is exact in FP32 (else FP64)
//   --lanczos            symmetric A: Lanczos three-term recurrence instead of Gram-Schmidt
//   --sstep S            s-step blocks of S basis vectors (2..16) after the first S columns
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../host/hip_check.hpp"
#include "mek/arnoldi_runner.hpp"
#include "mek/csr_host.hpp"

namespace {

bool ends_with(const std::string& s, const std::string& suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

bool has_bincsr_magic(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return false;
  unsigned char hdr[8] = {0};
  f.read(reinterpret_cast<char*>(hdr), 8);
  if (!f || f.gcount() != 8) return false;
  static const unsigned char magic[8] = {'C', 'S', 'R', 0, 0, 0, 0, 1};
  return std::memcmp(hdr, magic, 8) == 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  std::cout << "Starting matrix_exp application." << std::endl;
  bool use_file = false;
  std::string matrix_file;
  int test_size = 1000;
  int m_param = -1;
  int max_restarts_cli = MAX_RESTARTS;
  double t_cli = 1.0;
  double tol_cli = 1e-6;
  bool clamp_t = true;
  std::string v_file;
  mek::SolverOptions opts;

  std::vector<std::string> args(argv + 1, argv + argc);
  size_t pos_idx = 0;
  for (size_t i = 0; i < args.size(); ++i) {
    const std::string& a = args[i];
    const bool has_next = i + 1 < args.size();
    if (a == "--m" && has_next) { m_param = std::atoi(args[++i].c_str()); continue; }
    if (a == "--max-restarts" && has_next) { max_restarts_cli = std::atoi(args[++i].c_str()); continue; }
    if (a == "--t" && has_next) { t_cli = std::atof(args[++i].c_str()); continue; }
    if (a == "--tol" && has_next) { tol_cli = std::atof(args[++i].c_str()); continue; }
    if (a == "--no-t-clamp") { clamp_t = false; continue; }
    if (a == "--v-file" && has_next) { v_file = args[++i]; continue; }
    if (a == "--f32-values") { opts.values_f32 = true; continue; }
    if (a == "--lanczos") { opts.lanczos = true; continue; }
    if (a == "--sstep" && has_next) { opts.sstep = std::atoi(args[++i].c_str()); continue; }
    if (a == "--step-control" && has_next) {
      const std::string v = args[++i];
      opts.control = (v == "expokit")       ? mek::StepControl::Expokit
                     : (v == "expokit-rel") ? mek::StepControl::ExpokitRelative
                                            : mek::StepControl::Reference;
      continue;
    }
    if (a == "--gs" && has_next) {
      const std::string v = args[++i];
      opts.gs_mode = (v == "never") ? 0 : (v == "dgks") ? 2 : 1;
      continue;
    }
    // positional arguments (main.cu:2117-2140)
    if (pos_idx == 0) {
      if (a.find('.') != std::string::npos || a.find('/') != std::string::npos ||
          a.rfind(".mtx") != std::string::npos) {
        use_file = true;
        matrix_file = a;
      } else {
        test_size = std::atoi(a.c_str());
        if (test_size <= 0) {
          std::cerr << "Error: Matrix size must be a positive integer. Using default size 1000." << std::endl;
          test_size = 1000;
        }
      }
      ++pos_idx;
      continue;
    }
    if (pos_idx == 1) { m_param = std::atoi(a.c_str()); ++pos_idx; continue; }
    if (pos_idx == 2) { max_restarts_cli = std::atoi(a.c_str()); ++pos_idx; continue; }
  }

  try {
    int num_gpus = 0;
    MEK_HIP(hipGetDeviceCount(&num_gpus));
    std::cout << "Found " << num_gpus << " GPU(s)" << std::endl;
    if (num_gpus < 1) {
      std::cerr << "No GPUs found!" << std::endl;
      return EXIT_FAILURE;
    }
    CSRHost A;
    if (use_file) {
      std::cout << "Loading matrix from file: " << matrix_file << std::endl;
      const bool is_bincsr = ends_with(matrix_file, ".bincsr");
      const bool is_mtx = ends_with(matrix_file, ".mtx") || ends_with(matrix_file, ".mtx.gz");
      bool loaded = false;
      if (is_bincsr || has_bincsr_magic(matrix_file)) {
        A.from_binary_csr(matrix_file);
        loaded = true;
        std::cout << "Detected Binary CSR v1 input" << std::endl;
      } else if (is_mtx) {
        A.from_matrix_market(matrix_file);
        loaded = true;
        std::cout << "Detected MatrixMarket input" << std::endl;
      }
      if (!loaded) {
        std::cerr << "Error: Could not detect matrix format for file " << matrix_file
                  << ". Supported: .bincsr, .mtx, .mtx.gz" << std::endl;
        return EXIT_FAILURE;
      }
      std::cout << "Loaded matrix: " << A.rows << "x" << A.cols << " with " << A.nnz << " non-zeros"
                << std::endl;
    } else {
      // built-in diagonal test matrix a_ii = 1 + i/n (main.cu:2198-2216)
      std::cout << "Matrix size (generated): " << test_size << std::endl;
      A.rows = A.cols = test_size;
      A.nnz = test_size;
      A.row_ptr.resize(test_size + 1);
      A.col_idx.resize(test_size);
      A.values.resize(test_size);
      for (int i = 0; i < test_size; ++i) {
        A.row_ptr[i] = i;
        A.col_idx[i] = i;
        A.values[i] = 1.0 + static_cast<double>(i) / test_size;
      }
      A.row_ptr[test_size] = test_size;
      std::cout << "Created test matrix: " << A.rows << "x" << A.cols << " with " << A.nnz
                << " non-zeros" << std::endl;
    }
    const int num_gpus_to_use = 1;  // the solve runs on one GPU

    double max_abs = A.max_abs_value();
    if (max_abs <= 0.0) max_abs = 1.0;
    double t_safe = t_cli;
    const double t_limit = 10.0 / max_abs;
    if (clamp_t && t_safe > t_limit) {
      std::cout << "[WARN] Requested t=" << t_safe << " is large for max|A|=" << max_abs
                << ". Clamping t to " << t_limit << " to avoid overflow." << std::endl;
      t_safe = t_limit;
    }
    const int rows = A.rows, cols = A.cols;
    const long long nnz = A.nnz;
    // one GPU: the single row block takes A's arrays over (no copy)
    auto partitions = std::move(A).partition_rows(num_gpus_to_use);
    std::cout << "Partitioned matrix across " << num_gpus_to_use << " GPUs" << std::endl;
    std::vector<CSRHost::GhostMap> ghost_maps(num_gpus_to_use);
    for (int i = 0; i < num_gpus_to_use; ++i)
      ghost_maps[i] = partitions[i].build_owner_ghost_maps(partitions, i, num_gpus_to_use);
    validate_ghost_maps(ghost_maps, num_gpus_to_use);

    int m_val = m_param;
    if (m_val <= 0) m_val = std::min(30, std::max(5, rows - 1));
    m_val = std::max(1, std::min(m_val, rows - 1));
    ArnoldiParams params(m_val, t_safe, tol_cli, max_restarts_cli, true);
    std::cout << "Arnoldi params: m=" << params.m << " t=" << params.t << " tol=" << params.tol
              << " max_restarts=" << params.max_restarts << std::endl;

    ArnoldiRunner runner(num_gpus_to_use, params, rows, cols, nnz);
    runner.options = opts;
    runner.init_all_devices(partitions, {0}, ghost_maps);

    std::vector<double> v_host(rows, 0.1);
    if (!v_file.empty()) {
      std::ifstream vf(v_file);
      if (!vf) {
        std::cerr << "Error: could not open --v-file " << v_file << std::endl;
        return EXIT_FAILURE;
      }
      for (int i = 0; i < rows; ++i) {
        if (!(vf >> v_host[i])) {
          std::cerr << "Error: --v-file " << v_file << " holds fewer than " << rows << " values" << std::endl;
          return EXIT_FAILURE;
        }
      }
    }
    std::vector<double> y_host(rows, 0.0);
    runner.compute_expmv(v_host, y_host);

    if (const char* out_path = std::getenv("OUTPUT_Y_FILE"); out_path && out_path[0] != '\0') {
      std::ofstream ofs(out_path);
      if (!ofs) {
        std::cerr << "Error: could not open OUTPUT_Y_FILE=" << out_path << " for writing" << std::endl;
        return EXIT_FAILURE;
      }
      ofs.setf(std::ios::scientific);
      ofs.precision(16);
      for (double v : y_host) ofs << v << "\n";
      ofs.close();
      std::cout << "Saved y to " << out_path << std::endl;
    }
    std::cout << "Computation completed. Sample results:" << std::endl;
    for (int i = 0; i < std::min(10, static_cast<int>(y_host.size())); ++i)
      std::cout << "y[" << i << "] = " << y_host[i] << std::endl;
    std::cout << "Arnoldi summary: m=" << m_val << ", max_restarts=" << max_restarts_cli
              << ", restarts_done=" << runner.restarts_done
              << ", converged=" << (runner.converged ? "yes" : "no")
              << ", residual=" << runner.last_residual << std::endl;
    runner.destroy();
  } catch (const mek::Error& e) {
    std::cerr << e.what() << std::endl;
    return EXIT_FAILURE;
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return 0;
}
