// Device buffers and slots of the batched solver (batch.hip).
#pragma once

#include "common.hpp"
#include "mek/batch_runner.hpp"

namespace mek {

constexpr int kBatchNnzCap = 2048;  // nnz of a short-row SpMV block
constexpr int kBatchMaxM = 62;      // largest Krylov dimension of a batched problem
constexpr int kBatchRowsPerBlock = 4096;  // rows of one streaming block
constexpr int kBatchSweepGrid = 2048;     // workgroups of a per-column sweep (grid-stride over its list)
constexpr int kPScal = 16;          // per-problem double slots
constexpr int kPInt = 24;           // per-problem int slots

enum BatchScalar : int {
  kBVNorm = 0, kBWNorm2, kBWpNorm2, kBErr, kBYNorm, kBDt, kBTDone, kBLastRes,
  kBT, kBTol, kBDtMin, kBDtMax,
  kBAvNorm,   // Expokit: ||A v_{m+1}||
  kBTNew,     // Expokit: proposed next step
  kBSErr,     // Expokit: sum of local error estimates
  kBAnorm     // ||A||_inf (Expokit initial step and round-off floor)
};

enum BatchInt : int {
  kBActive = 0, kBNeedBasis, kBRestart, kBBreakdown, kBReorth, kBKUsed, kBM,
  kBAccepted, kBRejects, kBConverged, kBAborted, kBAttempts, kBColumns,
  kBReorthCount, kBMaxRestarts, kBMaxSteps, kBSubsteps,
  kBCtrl,      // 0 reference controller, 1 Expokit, 2 Expokit with relative local error
  kBIReject,   // Expokit: rejections of the current step
  kBLift,      // 1: lift y in this cycle
  kBKLift,     // basis vectors in the lift
  kBRejTotal   // all rejected attempts
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
  const int* p_rb0;         // P + 1: first SpMV row block of each problem
  const int* np_col;        // M: problems taking part in column j (a prefix: decreasing m)
  // Blocks of the problems that build a basis in this cycle, in storage
  // order (kb_compact at the start of every cycle): rb_list / sb_list, and
  // list_cnt[j] (j < M) row blocks of column j, list_cnt[M] all of them,
  // list_cnt[M + 1 + j] streaming blocks of column j, list_cnt[2M + 1] all.
  int* rb_list;
  int* sb_list;
  int* list_cnt;
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
  int* counters;            // [0] active problems, [1] cycles, [2] largest m of the active
  int any_expokit;          // 1 if some problem uses Expokit step control
};

void batch_prepare();
void batch_launch_init(const BatchBufs& b, hipStream_t s);
// The column-j SpMV over row blocks [0, nrb) with `grid` workgroups (every
// problem taken as active; timing and tuning at the start state).
void batch_launch_spmv(const BatchBufs& b, int j, int nrb, int grid, hipStream_t s);
// One cycle that builds mcols Arnoldi columns (the largest m of the problems
// still active; those with a smaller m skip the later columns).
void batch_launch_cycle(const BatchBufs& b, const BatchPlan& plan, int mcols, int gs_mode, double eta,
                        hipStream_t s);
// Largest exponential order of a cycle (M, or M + 2 with Expokit problems).
__host__ __device__ inline int batch_expm_order(const BatchBufs& b) { return b.any_expokit ? b.M + 2 : b.M; }

}  // namespace mek
