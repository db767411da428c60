// Device buffers and slots of the batched solver (batch.hip).
#pragma once

#include "common.hpp"

namespace mek {

constexpr int kBatchNnzCap = 2048;  // nnz of a short-row SpMV block
constexpr int kBatchMaxM = 62;      // largest Krylov dimension of a batched problem
constexpr int kBatchRowsPerBlock = 4096;  // rows of one streaming block
constexpr int kPScal = 16;          // per-problem double slots
constexpr int kPInt = 24;           // per-problem int slots

enum BatchScalar : int {
  kBVNorm = 0, kBWNorm2, kBWpNorm2, kBErr, kBYNorm, kBDt, kBTDone, kBLastRes,
  kBT, kBTol, kBDtMin, kBDtMax
};

enum BatchInt : int {
  kBActive = 0, kBNeedBasis, kBRestart, kBBreakdown, kBReorth, kBKUsed, kBM,
  kBAccepted, kBRejects, kBConverged, kBAborted, kBAttempts, kBColumns,
  kBReorthCount, kBMaxRestarts, kBMaxSteps, kBSubsteps
};

struct BatchBufs {
  int P;                    // problems
  int M;                    // largest m_p (basis columns allocated: M + 2)
  const int* row_ptr;       // block-diagonal CSR over all rows
  const int* col_idx;
  const void* vals;
  int vals_f32;
  const int* rb_start;      // SpMV row blocks (nrb + 1 starts), never across problems
  const int* rb_prob;
  int nrb;
  const int* sb_prob;       // streaming blocks: problem of each block
  const int* sb_row0;       // nsb + 1 row starts
  int nsb;
  const int* p_sb0;         // P + 1: first streaming block of each problem
  double* V;
  long long ld;
  double* w;
  double* y;
  double* colscale;         // P x (M+2)
  double* h1;               // P x (M+2)
  double* h2;               // P x (M+2)
  double* H;                // P x (M+2)^2
  double* wH;               // P x (M+4)
  double* pscal;            // P x kPScal
  int* pint;                // P x kPInt
  double* partials;         // (M+2) x nsb
  double* expm_ws;          // P x 6 M^2 when M > the LDS limit
  int* counters;            // [0] active problems, [1] cycles
};

void batch_prepare();
void batch_launch_init(const BatchBufs& b, hipStream_t s);
void batch_launch_cycle(const BatchBufs& b, int gs_mode, double eta, int spmv_grid, hipStream_t s);

}  // namespace mek
