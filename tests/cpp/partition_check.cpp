// Prints CSRHost::partition_rows and build_owner_ghost_maps of a Binary CSR
// file for the Python test to compare with its own restatement, and checks
// that an rvalue one-block partition takes the arrays over without a copy.
//
//   partition_check FILE P     partitions and ghost maps for P blocks
//   partition_check FILE move  the one-block move check
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "mek/csr_host.hpp"

namespace {

void print_list(const char* name, const std::vector<int>& v) {
  std::printf("%s", name);
  for (int x : v) std::printf(" %d", x);
  std::printf("\n");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  CSRHost A;
  A.load_binary_csr(argv[1]);
  if (std::strcmp(argv[2], "move") == 0) {
    const int* rp = A.row_ptr.data();
    const int* ci = A.col_idx.data();
    const double* va = A.values.data();
    const int rows = A.rows;
    const long long nnz = A.nnz;
    std::vector<CSRHost> parts = std::move(A).partition_rows(1);
    const bool same = parts.size() == 1 && parts[0].row_ptr.data() == rp && parts[0].col_idx.data() == ci &&
                      parts[0].values.data() == va && parts[0].rows == rows && parts[0].nnz == nnz;
    std::printf("moved %d emptied %d\n", same ? 1 : 0, A.col_idx.empty() && A.values.empty() ? 1 : 0);
    // a const matrix still gets copies
    const CSRHost& B = parts[0];
    std::vector<CSRHost> copies = B.partition_rows(1);
    std::printf("copied %d\n", copies[0].col_idx.data() != B.col_idx.data() && copies[0].col_idx == B.col_idx ? 1 : 0);
    return 0;
  }
  const int p = std::atoi(argv[2]);
  std::vector<CSRHost> parts = A.partition_rows(p);
  for (int i = 0; i < p; ++i) {
    const CSRHost& q = parts[i];
    std::printf("part %d rows %d cols %d nnz %lld\n", i, q.rows, q.cols, q.nnz);
    print_list("row_ptr", q.row_ptr);
    print_list("col_idx", q.col_idx);
    std::printf("values");
    for (double v : q.values) std::printf(" %.17g", v);
    std::printf("\n");
  }
  std::vector<CSRHost::GhostMap> maps(p);
  for (int i = 0; i < p; ++i) {
    maps[i] = A.build_owner_ghost_maps(parts, i, p);
    const CSRHost::GhostMap& g = maps[i];
    std::printf("map %d recv %zu send %zu\n", i, g.total_recv_size, g.total_send_size);
    for (const auto& kv : g.owner_to_consumer_map) {
      std::printf("recv_from %d", kv.first);
      for (int c : kv.second) std::printf(" %d", c);
      std::printf("\n");
    }
    for (const auto& kv : g.consumer_to_owner_map) {
      std::printf("send_to %d", kv.first);
      for (int c : kv.second) std::printf(" %d", c);
      std::printf("\n");
    }
    print_list("recv_offsets", g.recv_offsets);
    print_list("recv_counts", g.recv_counts);
    print_list("send_offsets", g.send_offsets);
    print_list("send_counts", g.send_counts);
    print_list("incoming", g.incoming_global_col_indices);
    print_list("outgoing", g.outgoing_global_col_indices);
  }
  validate_ghost_maps(maps, p);  // exits with a message on a mismatch
  std::printf("validated\n");
  return 0;
}
