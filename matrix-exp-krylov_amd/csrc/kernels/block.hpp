// Block Krylov exp(t A) V for a block of 32 right-hand sides on a shared A
// (block.hip): SpMM, block Gram-Schmidt on FP64 MFMA, small block algebra and
// a dense FP64 MFMA GEMM for the exponential of the block Hessenberg matrix.
#pragma once

#include "common.hpp"

namespace mek {

constexpr int kRhs = 32;              // right-hand sides per block
constexpr int kBlkRows = 4096;        // rows of one split-K chunk of the block dots
constexpr int kBlkMaxSteps = 48;      // largest number of block Arnoldi steps
constexpr int kGemmTile = 64;         // dense GEMM tile (square)

// Device buffers of one block solve.  Basis block i (n_pad x 32, row-major,
// rows >= n zero) is stored raw at P + i * n_pad * 32; its orthonormal form
// is Q_i = P_i S_i with the 32 x 32 upper-triangular S_i = R_i^{-1}.
struct BlockBufs {
  int n;                  // rows
  long long n_pad;        // rows rounded up to kBlkRows (zero padded)
  const int* row_ptr;     // CSR of A
  const int* col_idx;
  const double* vals;
  double* P;              // (m + 1) raw basis blocks
  double* W;              // n_pad x 32 work block
  double* Y;              // n_pad x 32 result block
  double* S;              // (m + 1) x 32 x 32 block scales (row-major)
  double* G;              // (m + 1) x 32 x 32 update coefficients
  double* C;              // (m + 1) x 32 x 32 reduced dot blocks
  double* Hb;             // block Hessenberg, ld = 32 (m + 1), column-major
  int ldh;
  double* R0;             // 32 x 32: V0 = Q_0 R0
  double* partials;       // split-K partials: [block][chunk][1024]
  int nchunks;            // n_pad / kBlkRows
  int* flags;             // [0] breakdown (rank-deficient block), [1] blocks built,
                          // [2] second pass needed for the current block, [3] second passes
  double* Gm;             // 32 x 32 Gram block of the new block from the first pass
  double* scal;           // [0] ||Y||_F, [1] residual estimate, [2] ||V||_F, [3] 1-norm
  int spmm_grid;          // workgroups of the SpMM and of the update sweep
  int update_grid;
};

// W = A * P_j (raw block j), rows < n; rows >= n of W are zero.
void block_launch_spmm(const BlockBufs& b, const double* X, double* Out, hipStream_t s);
// C_i = X1_i^T X2 for the nb consecutive blocks X1_i = X1 + i n_pad 32 and,
// with gram, C_nb = X2^T X2; reduced in fixed order into b.C[i].  cond: do
// nothing unless flags[2] asks for the second pass.
void block_launch_dots(const BlockBufs& b, const double* X1, int nb, const double* X2, int gram, int cond,
                       hipStream_t s);
// The two halves of block_launch_dots: the split-K partials (k_bdots) and
// their fixed-order reduction into b.C.
void block_launch_dots_partial(const BlockBufs& b, const double* X1, int nb, const double* X2, int gram,
                               int cond, hipStream_t s);
void block_launch_dots_fin(const BlockBufs& b, int nb, int gram, int cond, hipStream_t s);
// Out = X - sum_{i<nb} P_i G_i (X may be null: Out = -sum ...); cond as above.
void block_launch_update(const BlockBufs& b, int nb, const double* X, double* Out, int cond, hipStream_t s);
// One workgroup after the first pass of column j: the new block's Gram
// matrix by Pythagoras, Z^T Z - sum_i Hh_i^T Hh_i, and the second-pass test
// (gs_mode 0 never, 1 always, 2 when some column kept less than eta of its
// norm) into flags[2].
void block_launch_gram(const BlockBufs& b, int j, int gs_mode, double eta, hipStream_t s);
// Workgroup i < nb: Hh = S_i^T C_i into block (i, j) of H (pass 1) or added
// to it (pass 2, only when flags[2] asks for it), and the update
// coefficients G_i = S_i Hh.
void block_launch_coeffs(const BlockBufs& b, int j, int nb, int pass, hipStream_t s);
// One workgroup: Cholesky of the Gram block C[0] -> R, S_{j+1} = R^{-1}
// (j = -1: the start block, R0), H(j+1, j) = R S_j; breakdown flag.
void block_launch_chol(const BlockBufs& b, int j, hipStream_t s);
// G_i = S_i F_i for the lift (F: the first 32 columns of exp(dt H), times R0).
void block_launch_lift_coeffs(const BlockBufs& b, int nb, const double* F, int ldf, hipStream_t s);
// Residual estimate ||H(k, k-1) F_{k-1}||_F and ||F||_F into b.scal.
void block_launch_residual(const BlockBufs& b, int k, const double* F, int ldf, hipStream_t s);

// Dense FP64 MFMA GEMM, column-major, square tiles: D = alpha A B + beta E
// (E may alias D), N x N with N a multiple of kGemmTile.
void dense_gemm(int N, const double* A, int lda, const double* B, int ldb, double alpha,
                const double* E, int lde, double beta, double* D, int ldd, hipStream_t s);
// D = sum_k c_k X_k elementwise (up to 4 terms, null terms skipped), N x N.
void dense_combine(int N, int ld, double* D, const double* X0, double c0, const double* X1, double c1,
                   const double* X2, double c2, const double* X3, double c3, double diag, hipStream_t s);
// D (N x N) = scale * leading kk x kk block of src (ld lds), zero elsewhere.
void dense_load(int N, int kk, const double* src, int lds, double scale, double* D, hipStream_t s);
// 1-norm of an N x N matrix (max column abs sum) into *out (device).
void dense_norm1(int N, const double* A, int lda, double* out, hipStream_t s);

}  // namespace mek
