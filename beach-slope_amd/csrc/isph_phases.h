// isph_phases.h -- the particle phases of a production step for any number of
// independent scenarios, one launch per phase.
//
// Each scenario keeps its own arrays (its Dev); a launch over an index space
// (particles, cells, transfer bins or scan tiles) is the concatenation of
// every scenario's 256-wide blocks, and a workgroup finds its scenario by
// binary search over a prefix table of block counts (one table per index
// space).  The phase bodies are the single-scenario ones of isph_kernels.h,
// called with the scenario's own particle index, so a scenario's results do
// not depend on the others in the launch.  The scenario tables live in device
// memory: the ping-pong swap after a rebin is done there, in place, so a
// sequence of phases needs no host round trip.
#pragma once
#include "isph_kernels.h"

// index spaces of the prefix table
enum {
  PH_PART = 0,  // particles: the host's upper bound of the live count
  PH_CELL = 1,  // cells, plus one (cell starts)
  PH_BIN = 2,   // (cell, transfer class) bins, plus one
  PH_TILE = 3,  // scan tiles over the bins
  PH_P64 = 4,   // particles in blocks of 64 and of 128 threads (tuned launches)
  PH_P128 = 5,
  PH_T4 = 6,    // neighbour-search cell tiles of 4 and of 8 rows
  PH_T8 = 7,
  PH_NKIND = 8
};

struct PhaseBatch {
  int S;
  Dev* devs;            // S scenario tables (swapped in place by kb_swap)
  const int* pre;       // PH_NKIND x (S+1) block prefixes
  const double* wave;   // 3S: paddle X, U and 1 (0: no paddle) of the step
};

__device__ __forceinline__ bool phase_locate(const PhaseBatch& B, int kind, int gb, int& s,
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

__host__ __device__ inline int phase_nbins(const Dev& D) {
  return D.P.cellx * D.P.celly * NCLS + 1;
}

#define PH_IDX(B, kind)                                        \
  int s_, lb_;                                                 \
  if (!phase_locate(B, kind, blockIdx.x, s_, lb_)) return;     \
  const Dev& D = B.devs[s_];

// particle launches may use blocks of 64, 128 or 256 threads (the
// autotuner's choice); each block size has its own prefix table
__device__ __forceinline__ int ph_part_kind() {
  return blockDim.x == 64 ? PH_P64 : (blockDim.x == 128 ? PH_P128 : PH_PART);
}
#define PH_PIDX(B)                                             \
  PH_IDX(B, ph_part_kind());                                   \
  const int i = lb_ * (int)blockDim.x + threadIdx.x;           \
  const int n = D.sc->n;

// ping-pong swap of the gathered state (the rebin and buffer passes write
// the other half); host and device run the same swap
__host__ __device__ inline void dev_swap_state(Dev& D) {
#define SWP(a, b) { auto t_ = D.a; D.a = D.b; D.b = t_; }
  SWP(x, x2) SWP(y, y2) SWP(u, u2) SWP(v, v2) SWP(us, us2) SWP(vs, vs2) SWP(xs, xs2)
  SWP(ys, ys2) SWP(p, p2) SWP(m, m2) SWP(rho, rho2) SWP(rho0, rho02) SWP(vol0, vol02)
  SWP(c, c2) SWP(tid, tid2) SWP(flags, flags2) SWP(vic, vic2) SWP(cell, cell2)
#undef SWP
}

// ---- one thread per scenario ----
__global__ void kb_gather_phase(PhaseBatch B, DevScalars* out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < B.S) out[s] = *B.devs[s].sc;
}
__global__ void kb_advance_t(PhaseBatch B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < B.S) {
    DevScalars* S = B.devs[s].sc;
    S->t = S->t + S->dt;
  }
}
__global__ void kb_zero_nnz(PhaseBatch B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < B.S) B.devs[s].sc->nnz = 0;
}
__global__ void kb_swap(PhaseBatch B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < B.S) dev_swap_state(B.devs[s]);
}
// the error bits the phases raised (the step fails on them), cleared
__global__ void kb_clear_err(PhaseBatch B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < B.S) B.devs[s].sc->err = 0;
}

// ---- scalar transport: sub-step q of every scenario that has one ----
__global__ void __launch_bounds__(256) kb_scalart(PhaseBatch B, int q) {
  PH_PIDX(B);
  if (q >= D.sc->solsteps) return;
  scalart_body(D, i, n);
}
__global__ void kb_scalarupdate(PhaseBatch B, int q) {
  PH_PIDX(B);
  if (q >= D.sc->solsteps) return;
  scalarupdate_body(D, i, n, -1.0);
}

__global__ void kb_projection(PhaseBatch B) {
  PH_PIDX(B);
  projection_body(D, i, n);
}
__global__ void kb_paddle(PhaseBatch B) {
  PH_PIDX(B);
  if (B.wave[3 * s_ + 2] == 0.0) return;
  paddle_body(D, i, n, B.wave[3 * s_], B.wave[3 * s_ + 1]);
}

// ---- rebin (counting sort) ----
__global__ void kb_fill_bins(PhaseBatch B) {
  PH_IDX(B, PH_BIN);
  const int k = lb_ * 256 + threadIdx.x;
  if (k < phase_nbins(D)) D.bincnt[k] = 0;
}
__global__ void kb_rebin_classify(PhaseBatch B, int alt) {
  PH_PIDX(B);
  rebin_classify_body(D, i, n, alt);
}
__global__ void kb_rebin_rank(PhaseBatch B) {
  PH_PIDX(B);
  rebin_rank_body(D, i, n);
}
__global__ void kb_scan_tiles(PhaseBatch B) {
  PH_IDX(B, PH_TILE);
  scan_tile_body(D.bincnt, D.binstart, D.scan_tmp, phase_nbins(D), lb_);
}
__global__ void kb_scan_sums(PhaseBatch B) {
  const Dev& D = B.devs[blockIdx.x];
  scan_sums_body(D.scan_tmp, (phase_nbins(D) + SCAN_TILE - 1) / SCAN_TILE, nullptr);
}
__global__ void kb_scan_add(PhaseBatch B) {
  PH_IDX(B, PH_BIN);
  const int k = lb_ * 256 + threadIdx.x;
  if (k < phase_nbins(D)) D.binstart[k] += D.scan_tmp[k / SCAN_TILE];
}
__global__ void kb_rebin_scatter(PhaseBatch B) {
  PH_PIDX(B);
  rebin_scatter_body(D, i, n);
}
__global__ void kb_rebin_finish(PhaseBatch B) {
  PH_IDX(B, PH_CELL);
  rebin_finish_body(D, lb_ * 256 + threadIdx.x, D.P.cellx * D.P.celly);
}

// ---- neighbour-sum phases ----
__global__ void __launch_bounds__(256) kb_neighbour(PhaseBatch B) {
  PH_PIDX(B);
  neighbour_body(D, i, n);
}
__global__ void __launch_bounds__(128) kb_neighbour_tile4(PhaseBatch B) {
  PH_IDX(B, PH_T4);
  neighbour_tile<4, NT_CAP4>(D, lb_);
}
__global__ void __launch_bounds__(256) kb_neighbour_tile8(PhaseBatch B) {
  PH_IDX(B, PH_T8);
  neighbour_tile<8, NT_CAP8>(D, lb_);
}
__global__ void kb_effpor(PhaseBatch B) {
  PH_PIDX(B);
  effpor_body(D, i, n);
}
__global__ void kb_eddyvis(PhaseBatch B) {
  PH_PIDX(B);
  eddyvis_body(D, i, n);
}
__global__ void kb_int_vel_varts(PhaseBatch B) {
  PH_PIDX(B);
  int_vel_varts_body(D, i, n);
}
__global__ void __launch_bounds__(256) kb_int_vel(PhaseBatch B) {
  PH_PIDX(B);
  int_vel_body(D, i, n);
}
__global__ void __launch_bounds__(256) kb_ppe_assemble(PhaseBatch B) {
  PH_PIDX(B);
  ppe_assemble_body(D, i, n);
}
__global__ void __launch_bounds__(256) kb_comp_vel(PhaseBatch B) {
  PH_PIDX(B);
  comp_vel_body(D, i, n);
}
__global__ void __launch_bounds__(256) kb_cell_max(PhaseBatch B, int init) {
  PH_IDX(B, PH_CELL);
  cell_max_cell(D, lb_ * 256 + threadIdx.x, D.P.cellx * D.P.celly, init, lb_);
}
__global__ void kb_comp_pos(PhaseBatch B) {
  PH_PIDX(B);
  comp_pos_body(D, i, n);
}
__global__ void kb_freesurf(PhaseBatch B) {
  PH_PIDX(B);
  freesurf_body(D, i, n);
}
__global__ void __launch_bounds__(256) kb_opt2_shift(PhaseBatch B) {
  PH_PIDX(B);
  opt2_shift_body(D, i, n);
}
__global__ void kb_shift_apply(PhaseBatch B) {
  PH_PIDX(B);
  shift_apply_body(D, i, n);
}
__global__ void kb_densityupdate(PhaseBatch B) {
  PH_PIDX(B);
  densityupdate_body(D, i, n);
}
// one workgroup per scenario
__global__ void __launch_bounds__(256) kb_timestep(PhaseBatch B) {
  timestep_block(B.devs[blockIdx.x]);
}
