// Loads each file named on the command line with the throwing CSRHost
// loaders (Binary CSR v1 for *.bincsr, MatrixMarket otherwise) and prints
// one line per file: "ok rows cols nnz sum" or "error <message>".  Built by
// the tests with the host sanitizers around csr_host.cpp.
#include <cstdio>
#include <string>

#include "mek/csr_host.hpp"

int main(int argc, char** argv) {
  for (int i = 1; i < argc; ++i) {
    const std::string path = argv[i];
    CSRHost A;
    try {
      if (path.size() > 7 && path.compare(path.size() - 7, 7, ".bincsr") == 0) A.load_binary_csr(path);
      else A.load_matrix_market(path);
      double s = 0.0;
      for (long long k = 0; k < A.nnz; ++k) s += A.values[k] * (1.0 + A.col_idx[k]);
      for (int r = 0; r <= A.rows; ++r) s += A.row_ptr[r];
      std::printf("ok %d %d %lld %.17g\n", A.rows, A.cols, A.nnz, s);
    } catch (const mek::LoadError& e) {
      std::printf("error %s\n", e.what());
    }
  }
  return 0;
}
