// isph_batch.h -- kernels of the precision-form PPE solve for any number of
// independent scenarios: one launch per phase covers every scenario.
//
// Each scenario keeps its own arrays and scalars (its Dev and MgDev, copied
// into device tables before a solve).  A launch over an index space (the
// particle rows, or the cells of one multigrid level) is laid out as the
// concatenation of every scenario's 256-wide blocks; the prefix table
// pre[kind][s] holds the first global block of scenario s, and a workgroup
// finds its scenario by binary search (the table has S+1 entries, so about
// log2(S) cached loads).  Row launches also use the XCD-contiguous order of
// row_block(), so one XCD sweeps one contiguous run of rows across
// scenarios.  Finalize, coarse-level and dense kernels run one workgroup per
// scenario.  A scenario's results do not depend on which other scenarios
// share the launch: every sum is over its own rows, in its own block order.
#pragma once
#include "isph_mg.h"

// index spaces of the prefix table
enum {
  BK_ROWS = 0,                           // particle rows
  BK_CELL1 = 1,                          // level-1 cells
  BK_DOWN = 2,                           // + l-1: 4 x level-(l+1) aggregates, l <= nmb
  BK_UP = 2 + MG_MAX_LEVELS,             // + l-1: level-l cells (x 4 on level 1), l <= nmb
  BK_COARSEN = 2 + 2 * MG_MAX_LEVELS,    // + l-1: level-(l+1) cells, l < nlev
  BK_NKIND = 2 + 3 * MG_MAX_LEVELS
};

struct BatchDev {
  int S;
  const Dev* devs;      // S scenario tables
  const MgDev* mgs;     // S multigrid tables
  const int* pre;       // BK_NKIND x (S+1) block prefixes
  int rows;             // row blocks of all scenarios (the BK_ROWS total)
};

// scenario s and its local block lb of global block gb in index space kind
__device__ __forceinline__ bool batch_locate(const BatchDev& B, int kind, int gb, int& s,
                                             int& lb) {
  const int* p = B.pre + kind * (B.S + 1);
  if (gb >= p[B.S]) return false;
  int lo = 0, hi = B.S;  // p[lo] <= gb < p[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (p[mid] <= gb) lo = mid;
    else hi = mid;
  }
  s = lo;
  lb = gb - p[lo];
  return true;
}

// A batch of one (every single-context solve) skips the prefix table: its
// tables' addresses then follow from the kernel arguments alone, so the
// scenario's fields load in the second round of reads.  Row blocks past
// B.rows return; a cell launch's grid is exactly the table's total, so every
// cell workgroup is in range.
#define BATCH_ROW(B)                                                   \
  int s_ = 0, vb_ = row_block();                                       \
  if (vb_ >= B.rows) return;                                           \
  if (B.S > 1 && !batch_locate(B, BK_ROWS, vb_, s_, vb_)) return;      \
  const Dev& D = B.devs[s_];

#define BATCH_CELLS(B, kind)                                           \
  int s_ = 0, lb_ = blockIdx.x;                                        \
  if (B.S > 1 && !batch_locate(B, kind, blockIdx.x, s_, lb_)) return;  \
  const Dev& D = B.devs[s_];                                 \
  const MgDev& G = B.mgs[s_];                                \
  const int I_ = lb_ * 256 + threadIdx.x;                              \
  (void)D;

// ---- setup ----
__global__ void __launch_bounds__(256) kb_mg_setup0(BatchDev B) {
  BATCH_ROW(B);
  mg_setup0_row(D, B.mgs[s_], vb_ * RED_BLOCK + threadIdx.x);
}
__global__ void __launch_bounds__(256) kb_mg_stencil1(BatchDev B) {
  BATCH_CELLS(B, BK_CELL1);
  mg_stencil1_cell(D, G, I_);
}
__global__ void __launch_bounds__(256) kb_mg_coarsen(BatchDev B, int l) {
  BATCH_CELLS(B, BK_COARSEN + l - 1);
  mg_coarsen_cell(G, l, I_);
}
__global__ void __launch_bounds__(MG_CO_THREADS) kb_mg_dense(BatchDev B) {
  mg_dense_block(B.mgs[blockIdx.x]);
}

// ---- V-cycle: V & 1 = 0 applies M to p (pp32 -> r32), 1 to s (sv -> sv32);
// V & 2: the rows' packed 16-bit column offsets instead of the int32 columns
template <int V>
__global__ void __launch_bounds__(256) kb_mg_pre0(BatchDev B) {
  BATCH_ROW(B);
  const MgDev& G = B.mgs[s_];
  const int i = vb_ * RED_BLOCK + threadIdx.x;
  if ((V & 1) == 0) mg_pre0_row<(V & 2) != 0, float>(D, G, D.pp32, D.d32, i);
  else mg_pre0_row<(V & 2) != 0, double>(D, G, D.sv, D.d32, i);
}
__global__ void __launch_bounds__(256) kb_mg_restrict0(BatchDev B) {
  BATCH_CELLS(B, BK_CELL1);
  mg_restrict0_cell(D, G, D.d32, I_);
}
__global__ void __launch_bounds__(256) kb_mg_down(BatchDev B, int l) {
  BATCH_CELLS(B, BK_DOWN + l - 1);
  mg_lvl_down(G, l, I_);
}
__global__ void __launch_bounds__(MG_CO_THREADS) kb_mg_coarse(BatchDev B) {
  extern __shared__ double mg_lds[];
  mg_coarse_block(B.devs[blockIdx.x], B.mgs[blockIdx.x], mg_lds);
}
__global__ void __launch_bounds__(256) kb_mg_up(BatchDev B, int l) {
  BATCH_CELLS(B, BK_UP + l - 1);
  mg_lvl_up(D, G, l, I_);
}
template <int V>
__global__ void __launch_bounds__(256) kb_mg_post0(BatchDev B) {
  BATCH_ROW(B);
  const MgDev& G = B.mgs[s_];
  const int i = vb_ * RED_BLOCK + threadIdx.x;
  if ((V & 1) == 0) mg_post0_row<(V & 2) != 0, float>(D, G, D.pp32, D.r32, i);
  else mg_post0_row<(V & 2) != 0, double>(D, G, D.sv, D.sv32, i);
}

// ---- BiCGStab rows ----
template <bool OFF>
__global__ void __launch_bounds__(256) kb_pb_1(BatchDev B) {
  __shared__ double lds[RED_BLOCK];
  BATCH_ROW(B);
  pb_1_rows<OFF>(D, vb_, lds);
}
__global__ void __launch_bounds__(256) kb_pb_s(BatchDev B) {
  BATCH_ROW(B);
  pb_s_rows(D, B.mgs[s_], vb_);
}
template <bool OFF>
__global__ void __launch_bounds__(256) kb_pb_2(BatchDev B) {
  __shared__ double lds[5 * RED_BLOCK];
  BATCH_ROW(B);
  pb_2_rows<OFF>(D, vb_, lds);
}
__global__ void __launch_bounds__(256) kb_pb_34(BatchDev B, int seq) {
  BATCH_ROW(B);
  pb_34_rows(D, B.mgs[s_], vb_, seq);
}
__global__ void __launch_bounds__(256) kb_pb_resid(BatchDev B, int init) {
  __shared__ double lds[RED_BLOCK];
  BATCH_ROW(B);
  pb_resid_rows(D, B.mgs[s_], vb_, init, lds);
}
__global__ void __launch_bounds__(256) kb_pb_bnorm(BatchDev B) {
  __shared__ double lds[RED_BLOCK];
  BATCH_ROW(B);
  pb_bnorm_rows(D, vb_, lds);
}
// pressure scatter (isph.f90:468-486): the solution vector is indexed by row
__global__ void __launch_bounds__(256) kb_copy_sol_to_p(BatchDev B) {
  BATCH_ROW(B);
  const int i = vb_ * RED_BLOCK + threadIdx.x;
  if (i < D.sc->n) D.p[i] = D.sol[i];
  if (vb_ == 0 && threadIdx.x == 0) D.sc->far = 0;  // set again by the next setup if needed
}

// one workgroup per scenario
__global__ void __launch_bounds__(256) kb_pfinalize(BatchDev B, int op, int init, int seq) {
  __shared__ double lds[5 * RED_BLOCK];
  pb_finalize(B.devs[blockIdx.x], op, init, seq, lds);
}

// copy every scenario's scalars into one array (one host read per poll)
__global__ void kb_gather_scalars(BatchDev B, DevScalars* out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < B.S) out[s] = *B.devs[s].sc;
}
