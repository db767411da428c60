// CSRHost: host-side CSR matrix with the reference's loaders.
// reference: struct CSRHost, main.cu:86-692 (loaders :97-523,
// partition_rows :526-562, GhostMap/build_owner_ghost_maps :564-692).
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mek {

// Thrown by the non-exiting loader variants; what() is the reference's
// message text without the "Error: " prefix.
struct LoadError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

}  // namespace mek

struct CSRHost {
  int rows;
  int cols;
  long long nnz;
  std::vector<int> row_ptr;
  std::vector<int> col_idx;
  std::vector<double> values;
  // Additive: value dtype of the source file (1 = f32, 2 = f64; 2 for
  // MatrixMarket and generated matrices).  An f32 file widens to double
  // exactly, so storing it back in f32 on the device is lossless.
  int source_value_dtype = 2;

  CSRHost() : rows(0), cols(0), nnz(0) {}

  // Reference-compatible: on any error print "Error: <message>" to stderr and
  // exit(EXIT_FAILURE) (main.cu:97-245, main.cu:265-523).
  void from_matrix_market(const std::string& filename);
  void from_binary_csr(const std::string& filename);

  // Additive, library-friendly: same parsing and validation, throws
  // mek::LoadError instead of exiting.  MatrixMarket input may be gzip
  // compressed (.mtx.gz), which the reference accepts by name but cannot read.
  void load_matrix_market(const std::string& filename);
  void load_binary_csr(const std::string& filename);

  // reference: partition_rows (main.cu:526-562).  Balanced row blocks with
  // re-based row_ptr (copies of the arrays).
  std::vector<CSRHost> partition_rows(int num_gpus) const&;
  // Additive: the same on a matrix the caller gives up
  // (std::move(A).partition_rows(p)).  One block takes the arrays over
  // without a copy (A's arrays are left empty); for p > 1 the blocks are
  // copies, as above.
  std::vector<CSRHost> partition_rows(int num_gpus) &&;

  // reference: GhostMap (main.cu:567-585).  Kept for API compatibility; on the
  // single-GPU path every map is empty.
  struct GhostMap {
    std::vector<int> ghost_indices;
    std::vector<int> ghost_owners;
    std::map<int, std::vector<int>> owner_to_consumer_map;
    std::map<int, std::vector<int>> consumer_to_owner_map;
    std::vector<int> send_offsets;
    std::vector<int> recv_offsets;
    std::vector<int> send_counts;
    std::vector<int> recv_counts;
    std::vector<int> outgoing_global_col_indices;
    std::vector<int> incoming_global_col_indices;
    std::map<int, std::pair<size_t, size_t>> outgoing_peer_info;
    std::map<int, std::pair<size_t, size_t>> incoming_peer_info;
    size_t total_recv_size = 0;
    size_t total_send_size = 0;
  };

  // reference: build_owner_ghost_maps (main.cu:587-692): for partition
  // `my_gpu_id`, the columns it needs from other row blocks (receive lists)
  // and the columns others need from it (send lists), sorted and unique.
  GhostMap build_owner_ghost_maps(const std::vector<CSRHost>& partitions, int my_gpu_id,
                                  int num_gpus) const;

  // Additive helpers.
  double max_abs_value() const;       // max |a_ij| (0 for an empty matrix)
  double norm_inf() const;            // max_i sum_j |a_ij|
  double norm_1() const;              // max_j sum_i |a_ij| (sums in storage order)
  // Values that FP32 storage would change: their count, and the largest
  // relative change |fl32(a) - a| / |a| among the finite ones in *max_rel (0
  // when all are exact).  Values that are not finite, or whose FP32 image is
  // not (beyond FP32's range), are counted and also reported in *nonfinite.
  long long f32_inexact(double* max_rel, long long* nonfinite = nullptr) const;
};

// reference: validate_ghost_maps (main.cu:695-744).  Prints the mismatch and
// exits, like the reference.
void validate_ghost_maps(const std::vector<CSRHost::GhostMap>& ghost_maps, int num_gpus);
