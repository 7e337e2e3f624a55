// isph_kernels.h -- HIP kernels of the per-step ISPH pipeline (gfx950).
//
// One thread per particle (or PPE row) unless stated; blocks of 256 threads
// (four wave64s).  Kernels read the live particle count from the device
// scalars, so launch shapes depend only on a host-side upper bound and a
// step can be replayed without host round trips.
#pragma once
#include "isph_device.h"

#define TID_FIXED 1
#define TID_GHOST 2
#define TID_FLUID 3
#define F_BUFFER 1
#define F_FREE 2
#define F_PADDLE 8

#define PIDX()                                   \
  const int i = blockIdx.x * blockDim.x + threadIdx.x; \
  const int n = D.sc->n;

__global__ void k_fill_i32(int32_t* a, int len, int val) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) a[i] = val;
}

// ---------------------------------------------------------------------
// projection, isph.f90:17-47
// ---------------------------------------------------------------------
__device__ __forceinline__ void projection_body(const Dev& D, const int i, const int n) {
  if (i >= n || D.tid[i] != TID_FLUID) return;
  const double dt = D.sc->dt;
  D.xs[i] = D.x[i];
  D.x[i] = D.x[i] + ((dt * D.u[i]) / D.por[i]);
  D.ys[i] = D.y[i];
  D.y[i] = D.y[i] + ((dt * D.v[i]) / D.por[i]);
}
__global__ void k_projection(Dev D) {
  PIDX();
  projection_body(D, i, n);
}

// ---------------------------------------------------------------------
// piston wavemaker (no counterpart in the reference): the paddle's wall
// particles sit at rest position xs + X and move with velocity U along x.
// The PPE source term reads a wall neighbour's intermediate velocity (us, vs),
// which int_vel leaves to walls, so the paddle sets it too.
// ---------------------------------------------------------------------
__device__ __forceinline__ void paddle_body(const Dev& D, const int i, const int n, double X, double U) {
  if (i >= n || D.tid[i] == TID_FLUID || !(D.flags[i] & F_PADDLE)) return;
  D.x[i] = D.xs[i] + X;
  D.u[i] = U;
  D.v[i] = 0.0;
  D.us[i] = U;
  D.vs[i] = 0.0;
}
__global__ void k_paddle(Dev D, double X, double U) {
  PIDX();
  paddle_body(D, i, n, X, U);
}
__device__ __forceinline__ void paddle_rest_body(const Dev& D, const int i, const int n, double X) {
  if (i >= n || D.tid[i] == TID_FLUID || !(D.flags[i] & F_PADDLE)) return;
  D.xs[i] = D.x[i] - X;
}
__global__ void k_paddle_rest(Dev D, double X) {
  PIDX();
  paddle_rest_body(D, i, n, X);
}

// ---------------------------------------------------------------------
// rebin = cellshift / cellshiftalt (integrator.f90:33-392).
// A particle moves at most one cell per call, so its destination is fixed by
// its transfer class: 0 wall, 1 stay, then tn ts tw te tse tne tsw tnw in the
// order the receiving cell appends them (integrator.f90:324-386).  Sorting by
// the (destination cell, class) bin, stably in the current order, gives the
// reference's slot order exactly: one counting-sort pass with prefix-summed
// bin starts (the bins of a cell are consecutive, so its first bin start is
// the cell start).  Wavemaker paddle walls take classes 1-9 by position like
// fluid, so they follow the fixed walls of their cell.
// ---------------------------------------------------------------------
__device__ __forceinline__ int dest_offset_ci(int cls) {
  // row offset of the destination cell relative to the source cell
  switch (cls) {
    case 2: return -1;  // tn: north = up = row - 1
    case 3: return +1;  // ts
    case 6: return +1;  // tse
    case 7: return -1;  // tne
    case 8: return +1;  // tsw
    case 9: return -1;  // tnw
    default: return 0;
  }
}
__device__ __forceinline__ int dest_offset_cj(int cls) {
  switch (cls) {
    case 4: return -1;  // tw
    case 5: return +1;  // te
    case 6: return +1;  // tse
    case 7: return +1;  // tne
    case 8: return -1;  // tsw
    case 9: return -1;  // tnw
    default: return 0;
  }
}

__device__ __forceinline__ int rebin_bin(const Dev& D, int i, int alt) {
  const DevParams& P = D.P;
  int cls;
  const int key = D.cell[i];
  // walls stay in their slots; paddle walls move between cells as fluid does
  if (D.tid[i] != TID_FLUID && !(D.flags[i] & F_PADDLE)) {
    cls = 0;
  } else {
    int ci, cj;
    double xl, xr, yb, yt;
    cell_bounds(key, P, ci, cj, xl, xr, yb, yt);
    const double x1 = D.x[i], y1 = D.y[i];
    cls = CLS_DROP;
    if ((x1 >= xl) && (x1 < xr)) {
      if ((y1 >= yb) && (y1 < yt)) cls = 1;
      else if (y1 < yb) cls = 3;
      else if (y1 >= yt) cls = 2;
    } else if (x1 >= xr) {
      if (y1 >= yt) cls = 7;
      else if ((y1 >= yb) && (y1 < yt)) cls = 5;
      else if (y1 < yb) cls = 6;
    } else if (x1 < xl) {
      if (y1 >= yt) cls = 9;
      else if ((y1 >= yb) && (y1 < yt)) cls = 4;
      else if (y1 < yb) cls = 8;
    }
    if (alt && cls != CLS_DROP && D.tid[i] == TID_FLUID) {
      const bool buf = (D.flags[i] & F_BUFFER) != 0;
      if (!((x1 > P.xlcutoff) || buf || ((x1 <= P.xlcutoff) && (y1 <= P.ytcutoff)))) cls = CLS_DROP;
    }
    if (cls != CLS_DROP) {
      const int dci = ci + dest_offset_ci(cls), dcj = cj + dest_offset_cj(cls);
      // only interior cells are rebuilt (integrator.f90:320-321)
      if (dci < 1 || dci > P.celly - 2 || dcj < 1 || dcj > P.cellx - 2) cls = CLS_DROP;
    }
  }
  if (cls == CLS_DROP) return -1;
  const int ci = key % P.celly, cj = key / P.celly;
  const int dkey = (cj + dest_offset_cj(cls)) * P.celly + (ci + dest_offset_ci(cls));
  return dkey * NCLS + cls;
}

// The lanes of a wavefront hold consecutive particles, so a run of lanes with
// one bin (a cell's stayers, mostly) adds its length with one atomic from its
// first lane.  The counts are the same integers as one atomic per particle.
__device__ __forceinline__ void rebin_classify_body(const Dev& D, const int i, const int n, int alt) {
  const int bin = (i < n) ? rebin_bin(D, i, alt) : -2;
  if (i < n) D.pbin[i] = bin;
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(bin, 1, 64);
  const bool head = (lane == 0) || (prev != bin);
  const unsigned long long heads = __ballot(head);
  const unsigned long long later = (lane == 63) ? 0ull : (heads >> (lane + 1));
  const int len = later ? __ffsll((long long)later) : 64 - lane;
  if (head && bin >= 0) atomicAdd(&D.bincnt[bin], len);
}
__global__ void k_rebin_classify(Dev D, int alt) {
  PIDX();
  rebin_classify_body(D, i, n, alt);
}

// rank of a particle among the earlier particles of its source cell with the
// same bin (all particles of one bin come from one source cell)
__device__ __forceinline__ void rebin_rank_body(const Dev& D, const int i, const int n) {
  if (i >= n) return;
  const int bin = D.pbin[i];
  if (bin < 0) return;
  const int s0 = D.cstart[D.cell[i]];
  int rank = 0;
  for (int j0 = s0; j0 < i; j0 += 8) {  // eight loads in flight
    int b[8];
#pragma unroll
    for (int u = 0; u < 8; u++) b[u] = (j0 + u < i) ? D.pbin[j0 + u] : -1;
#pragma unroll
    for (int u = 0; u < 8; u++) rank += (j0 + u < i && b[u] == bin);
  }
  D.prank[i] = rank;
}
__global__ void k_rebin_rank(Dev D) {
  PIDX();
  rebin_rank_body(D, i, n);
}

// scatter every gathered field into the new order (ping-pong buffers)
__device__ __forceinline__ void rebin_scatter_body(const Dev& D, const int i, const int n) {
  if (i >= n) return;
  const int bin = D.pbin[i];
  if (bin < 0) return;
  const int d = D.binstart[bin] + D.prank[i];
  D.x2[d] = D.x[i]; D.y2[d] = D.y[i]; D.u2[d] = D.u[i]; D.v2[d] = D.v[i];
  D.us2[d] = D.us[i]; D.vs2[d] = D.vs[i]; D.xs2[d] = D.xs[i]; D.ys2[d] = D.ys[i];
  D.p2[d] = D.p[i]; D.m2[d] = D.m[i]; D.rho2[d] = D.rho[i]; D.rho02[d] = D.rho0[i];
  D.vol02[d] = D.vol0[i]; D.c2[d] = D.c[i];
  D.tid2[d] = D.tid[i]; D.flags2[d] = D.flags[i]; D.vic2[d] = D.vic[i];
  D.cell2[d] = bin / NCLS;
}
__global__ void k_rebin_scatter(Dev D) {
  PIDX();
  rebin_scatter_body(D, i, n);
}

__device__ __forceinline__ void rebin_finish_body(const Dev& D, int c, int ncell) {
  if (c <= ncell) D.cstart[c] = D.binstart[c * NCLS];
  if (c == 0) D.sc->n = D.binstart[ncell * NCLS];
}
__global__ void k_rebin_finish(Dev D, int ncell) {
  rebin_finish_body(D, blockIdx.x * blockDim.x + threadIdx.x, ncell);
}

// ---------------------------------------------------------------------
// exclusive scan of int32 (bins -> bin starts), three passes
// ---------------------------------------------------------------------
#define SCAN_ITEMS 4
#define SCAN_TILE (256 * SCAN_ITEMS)
__device__ __forceinline__ void scan_tile_body(const int32_t* in, int32_t* out,
                                               int32_t* tile_sums, int len, int tile) {
  __shared__ int32_t s[256];
  const int t = threadIdx.x;
  const int base = tile * SCAN_TILE + t * SCAN_ITEMS;
  int32_t v[SCAN_ITEMS], sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = (base + k < len) ? in[base + k] : 0;
    sum += v[k];
  }
  s[t] = sum;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int32_t a = (t >= off) ? s[t - off] : 0;
    __syncthreads();
    s[t] += a;
    __syncthreads();
  }
  int32_t run = s[t] - sum;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < len) out[base + k] = run;
    run += v[k];
  }
  if (t == 255) tile_sums[tile] = s[255];
}
__global__ void k_scan_tiles(const int32_t* in, int32_t* out, int32_t* tile_sums, int len) {
  scan_tile_body(in, out, tile_sums, len, blockIdx.x);
}

__device__ __forceinline__ void scan_sums_body(int32_t* tile_sums, int ntiles, int32_t* total_out) {
  __shared__ int32_t s[256];
  const int t = threadIdx.x;
  int32_t carry = 0;
  for (int base = 0; base < ntiles; base += 256) {
    int32_t v = (base + t < ntiles) ? tile_sums[base + t] : 0;
    s[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      int32_t a = (t >= off) ? s[t - off] : 0;
      __syncthreads();
      s[t] += a;
      __syncthreads();
    }
    if (base + t < ntiles) tile_sums[base + t] = carry + s[t] - v;
    int32_t tot = s[255];
    __syncthreads();
    carry += tot;
  }
  if (t == 0 && total_out) *total_out = carry;
}
__global__ void k_scan_sums(int32_t* tile_sums, int ntiles, int32_t* total_out) {
  scan_sums_body(tile_sums, ntiles, total_out);
}

__global__ void k_scan_add(int32_t* out, const int32_t* tile_sums, int len) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < len) out[idx] += tile_sums[idx / SCAN_TILE];
}

// ---------------------------------------------------------------------
// neighbour / nsearch2, search.f90:13-399
// ---------------------------------------------------------------------
// The search for particle i.  POS(j, x, y) gives particle j's position,
// MR(j, m, rho) its mass and density: read from global memory by the
// per-particle kernel, from the workgroup's LDS tile by the tiled one; the
// arithmetic and its order are the same, so both give the same bits.
template <typename POS, typename MR>
__device__ __forceinline__ void neighbour_core(const Dev& D, const int i, POS pos, MR mr) {
  const DevParams& P = D.P;
  double xi, yi;
  pos(i, xi, yi);
  int ci, cj;
  double xl, xr, yb, yt;
  cell_bounds(D.cell[i], P, ci, cj, xl, xr, yb, yt);
  const int q = quadrant(xi, yi, xl, xr, yb, yt);
  const int r0 = ci - ((q == 2 || q == 3) ? 1 : 0);
  const int c0 = cj - ((q == 0 || q == 3) ? 1 : 0);
  int cnt = 0;
  const int cap = D.cap;
  // Two passes, so the FP64 work of the accepted pairs runs with the lanes of
  // a wavefront together rather than under the candidate test's divergence:
  // the candidates are tested and the list written first, then the list is
  // walked in its order.  Each pass requests a batch of positions (NB_CAND
  // candidates, NB_PAIR neighbours) before testing or using any of them.  A
  // lane walks ~66 candidates of four cells; batches of 8 candidates took
  // 8% off the kernel, and larger batches of either kind gained nothing
  // more while costing registers.  r is recomputed from the same operands,
  // so every value and sum is the one of a single pass.
  constexpr int NB_CAND = 8, NB_PAIR = 2;
#pragma unroll 1
  for (int half = 0; half < 2; half++) {
    const int k0 = (c0 + half) * P.celly + r0;
    const int jb = D.cstart[k0], je = D.cstart[k0 + 2];
    for (int b0 = jb; b0 < je; b0 += NB_CAND) {
      double cx[NB_CAND], cy[NB_CAND];
#pragma unroll
      for (int u = 0; u < NB_CAND; u++) {
        const int j = (b0 + u < je) ? b0 + u : b0;
        pos(j, cx[u], cy[u]);
      }
#pragma unroll
      for (int u = 0; u < NB_CAND; u++) {
        const int j = b0 + u;
        if (j >= je) break;
        if (j == i) continue;
        const double dx = xi - cx[u], dy = yi - cy[u];
        const double d2 = (dx * dx) + (dy * dy);
        if (d2 > P.lim2) continue;
        if (!(sqrt(d2) < P.two_h)) continue;
        if (cnt < NBR_CAP) D.nbr[cnt * cap + i] = j;
        cnt++;
      }
    }
  }
  if (cnt > NBR_CAP) {
    // the step fails with ISPH_ECAP; the sums below cover the stored entries
    atomicOr(&D.sc->err, ERR_NBR_CAP);
    cnt = NBR_CAP;
  }
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
  for (int m0k = 0; m0k < cnt; m0k += NB_PAIR) {
    double xb[NB_PAIR], yb[NB_PAIR], mb[NB_PAIR], rb[NB_PAIR];
    int jn[NB_PAIR];
#pragma unroll
    for (int q = 0; q < NB_PAIR; q++) jn[q] = (m0k + q < cnt) ? D.nbr[(m0k + q) * cap + i] : i;
#pragma unroll
    for (int q = 0; q < NB_PAIR; q++) {
      const int j = jn[q];
      pos(j, xb[q], yb[q]);
      mr(j, mb[q], rb[q]);
    }
#pragma unroll
    for (int q = 0; q < NB_PAIR; q++) {
      if (m0k + q >= cnt) break;
      const double xj = xb[q], yj = yb[q];
      const double dx = xi - xj, dy = yi - yj;
      const double r = sqrt((dx * dx) + (dy * dy));
      double k1, k2;
      d_wabd2(dx, dy, r, P, k1, k2);
      const double mj = mb[q], rj = rb[q];
      const double r1 = (mj * (xj - xi)) / rj;
      const double r3 = (mj * (yj - yi)) / rj;
      m0 = m0 + k1 * r1;
      m1 = m1 + k2 * r1;
      m2 = m2 + k1 * r3;
      m3 = m3 + k2 * r3;
    }
  }
  D.ncnt[i] = cnt;
  D.xnb[i] = xi;
  D.ynb[i] = yi;
  uint8_t fl = D.flags[i] & (uint8_t)~F_FREE;
  double gv = 0.0;
  if (D.tid[i] == TID_FLUID) {
    D.vic[i] = 0;
    gv = m0 + m3;
    // invertmat2D, functions.f90:62-96 (1e-5 is a real4 literal)
    const double det = m0 * m3 - m1 * m2;
    double c1, c2, c3, c4;
    if ((fabs(det) >= (double)1e-5f) && !isnan(det) && isfinite(det)) {
      c1 = m3 / det; c2 = (-m1) / det; c3 = (-m2) / det; c4 = m0 / det;
    } else {
      c1 = 1.0; c2 = 0.0; c3 = 0.0; c4 = 1.0;
    }
    D.cf1[i] = c1; D.cf2[i] = c2; D.cf3[i] = c3; D.cf4[i] = c4;
    if ((gv <= P.fs_thresh) && !(fl & F_BUFFER)) fl |= F_FREE;
  }
  D.gradvx[i] = gv;
  D.flags[i] = fl;
}

__device__ __forceinline__ void neighbour_body(const Dev& D, const int i, const int n) {
  if (i >= n) return;
  neighbour_core(
      D, i, [&](int j, double& x, double& y) { x = D.x[j]; y = D.y[j]; },
      [&](int j, double& m, double& r) { m = D.m[j]; r = D.rho[j]; });
}
__global__ void __launch_bounds__(256) k_neighbour(Dev D) {
  PIDX();
  neighbour_body(D, i, n);
}

// ---------------------------------------------------------------------
// The neighbour search with its cells staged in LDS.  A workgroup of
// 32 ROWS threads takes a tile of ROWS cells of one cell column.  A particle
// of cell (row ci, column cj) searches the 2 x 2 cells of its quadrant:
// columns cj-1..cj+1 and rows ci-1..ci+1 hold every candidate, so the tile's
// particles find theirs in the three columns around it, rows one below to
// one above the tile, and each column's part of that is one contiguous range
// of particle indices (cell-sorted order).  The workgroup loads the three
// ranges' positions, masses and densities into LDS with coalesced loads,
// then each thread searches for particles of the tile reading LDS only.  A
// tile whose neighbourhood exceeds CAP particles, or that holds a particle in
// an edge cell of the grid, searches from global memory instead.  Lists, sums and flags are
// those of the per-particle kernel, bit for bit (the same core, other
// loaders).
// ---------------------------------------------------------------------
__host__ __device__ inline int ntile_count(int rows, int cellx, int celly) {
  return cellx * ((celly + rows - 1) / rows);
}

template <int ROWS, int CAP>
__device__ __forceinline__ void neighbour_tile(const Dev& D, const int tile) {
  __shared__ double lx[CAP], ly[CAP], lm[CAP], lr[CAP];
  const int cellx = D.P.cellx, celly = D.P.celly;
  const int tpc = (celly + ROWS - 1) / ROWS;  // tiles per column
  const int cj = tile / tpc, r0 = (tile - cj * tpc) * ROWS;
  if (cj >= cellx) return;                   // the whole workgroup
  const int r1 = min(r0 + ROWS, celly);      // tile rows [r0, r1)
  const int p0 = D.cstart[cj * celly + r0], p1 = D.cstart[cj * celly + r1];
  if (p1 <= p0) return;
  const int n = D.sc->n;
  // a particle in an edge cell of the grid would search past it: such tiles
  // search from global memory
  const int k = cj * celly;
  if (cj < 1 || cj + 1 >= cellx || (r0 == 0 && D.cstart[k + 1] > D.cstart[k]) ||
      (r1 == celly && D.cstart[k + celly] > D.cstart[k + celly - 1])) {
    for (int i = p0 + (int)threadIdx.x; i < p1; i += blockDim.x) neighbour_body(D, i, n);
    return;
  }
  // the three columns' ranges over rows r0-1 .. r1 (within the grid)
  const int lo = max(r0 - 1, 0), hi = min(r1 + 1, celly);
  int gs[3], ge[3], off[3];
  int tot = 0;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int kq = (cj - 1 + q) * celly;
    gs[q] = D.cstart[kq + lo];
    ge[q] = D.cstart[kq + hi];
    off[q] = tot;
    tot += ge[q] - gs[q];
  }
  if (tot > CAP) {
    for (int i = p0 + (int)threadIdx.x; i < p1; i += blockDim.x) neighbour_body(D, i, n);
    return;
  }
#pragma unroll
  for (int q = 0; q < 3; q++)
    for (int t = threadIdx.x; t < ge[q] - gs[q]; t += blockDim.x) {
      const int j = gs[q] + t, l = off[q] + t;
      lx[l] = D.x[j]; ly[l] = D.y[j]; lm[l] = D.m[j]; lr[l] = D.rho[j];
    }
  __syncthreads();
  // particle index -> LDS slot (the ranges are ordered by column)
  auto slot = [&](int j) {
    return j >= gs[2] ? off[2] + (j - gs[2]) : (j >= gs[1] ? off[1] + (j - gs[1]) : j - gs[0]);
  };
  for (int i = p0 + (int)threadIdx.x; i < p1; i += blockDim.x)
    neighbour_core(
        D, i, [&](int j, double& x, double& y) { const int l = slot(j); x = lx[l]; y = ly[l]; },
        [&](int j, double& m, double& r) { const int l = slot(j); m = lm[l]; r = lr[l]; });
}
// the two tile shapes: 4 rows on 128 threads, 8 rows on 256
#define NT_CAP4 640
#define NT_CAP8 1152
__global__ void __launch_bounds__(128) k_neighbour_tile4(Dev D) {
  neighbour_tile<4, NT_CAP4>(D, blockIdx.x);
}
__global__ void __launch_bounds__(256) k_neighbour_tile8(Dev D) {
  neighbour_tile<8, NT_CAP8>(D, blockIdx.x);
}

// ---------------------------------------------------------------------
// effpor + porsearch, porous.f90:32-224
// ---------------------------------------------------------------------
__device__ __forceinline__ void effpor_body(const Dev& D, const int i, const int n) {
  if (i >= n) return;
  const DevParams& P = D.P;
  D.resist[i] = 0.0;
  if (D.tid[i] != TID_FLUID) {
    D.por[i] = 1.0;  // non-fluid porosity keeps its default (particle.f90:62-64)
    return;
  }
  int ci, cj;
  double xl, xr, yb, yt;
  cell_bounds(D.cell[i], P, ci, cj, xl, xr, yb, yt);
  const double x = D.x[i], y = D.y[i];
  const int q = quadrant(x, y, xl, xr, yb, yt);
  // self first, then the quadrant's other three cells (porous.f90:138-176)
  int di[4] = {0, 0, 0, 0}, dj[4] = {0, 0, 0, 0};
  if (q == 0) { dj[1] = -1; di[2] = 1; di[3] = 1; dj[3] = -1; }
  else if (q == 1) { dj[1] = 1; di[2] = 1; di[3] = 1; dj[3] = 1; }
  else if (q == 2) { di[1] = -1; dj[1] = 1; di[2] = -1; dj[3] = 1; }
  else { di[1] = -1; di[2] = -1; dj[2] = -1; dj[3] = -1; }
  const double por = P.por;
  double sf = 0.0;
  // eight solids' positions requested before any is tested, then the mass
  // and density of those within reach requested before any is used (the sum
  // stays in solid order)
  constexpr int NB_SOL = 8;
  for (int cb = 0; cb < 4; cb++) {
    const int k = (cj + dj[cb]) * P.celly + (ci + di[cb]);
    const int qe = D.sstart[k + 1];
    for (int s0 = D.sstart[k]; s0 < qe; s0 += NB_SOL) {
      double sx[NB_SOL], sy[NB_SOL], d2[NB_SOL], sm[NB_SOL], sr[NB_SOL];
#pragma unroll
      for (int u = 0; u < NB_SOL; u++) {
        const int s = (s0 + u < qe) ? s0 + u : s0;
        sx[u] = D.sx[s];
        sy[u] = D.sy[s];
      }
#pragma unroll
      for (int u = 0; u < NB_SOL; u++) {
        const double dx = x - sx[u], dy = y - sy[u];
        d2[u] = (dx * dx) + (dy * dy);
        // certainly beyond 2h (or past the cell's solids): nothing to read
        const bool in = (s0 + u < qe) && !(d2[u] > P.lim2);
        const int s = in ? s0 + u : s0;
        sm[u] = in ? D.sm[s] : 0.0;
        sr[u] = in ? D.srho[s] : 1.0;
      }
#pragma unroll
      for (int u = 0; u < NB_SOL; u++) {
        if (s0 + u >= qe) break;
        if (d2[u] > P.lim2) continue;  // no square root
        const double r = sqrt(d2[u]);
        if (r < P.two_h) sf = sf + ((((1.0 - por * 1.0) * d_wab(r, P)) * sm[u]) / sr[u]);
      }
    }
  }
  if (sf > (1.0 - por)) sf = 1.0 - por;
  const double nn = 1.0 - sf;
  D.por[i] = nn;
  D.rho[i] = nn * D.rho0[i];
  const double n3 = (nn * nn) * nn;
  const double fch = 1.750 / sqrt(150 * n3);
  const double ui = D.u[i], vi = D.v[i];
  const double sp = sqrt((ui * ui) + (vi * vi));
  const double a = -((((d_vis(D.rho0[i]) / D.rho0[i]) * P.lamk) * ((1.0 - nn) * (1.0 - nn))) /
                     (n3 * (P.d50 * P.d50)));
  const double b = (((fch * sp) * sqrt(P.lamk)) * (1.0 - nn)) / (sqrt(n3) * P.d50);
  D.resist[i] = a - b;
}
__global__ void k_effpor(Dev D) {
  PIDX();
  effpor_body(D, i, n);
}

// ---------------------------------------------------------------------
// eddyvis, turbulence.f90:12-110
// ---------------------------------------------------------------------
__device__ __forceinline__ void eddyvis_body(const Dev& D, const int i, const int n) {
  if (i >= n) return;
  if (D.tid[i] != TID_FLUID) {
    D.nut[i] = 0.0;
    D.tke[i] = 0.0;
    return;
  }
  const DevParams& P = D.P;
  const int cap = D.cap, cnt = D.ncnt[i];
  const double c1 = D.cf1[i], c2 = D.cf2[i], c3 = D.cf3[i], c4 = D.cf4[i];
  const double xi = D.xnb[i], yi = D.ynb[i];
  const double ui = D.u[i] / D.por[i], vi = D.v[i] / D.por[i];
  double t1 = 0.0, t2 = 0.0, t3 = 0.0;
  // neighbours two at a time: both neighbours' fields are requested before
  // either is used (the sums stay in list order)
  for (int k0 = 0; k0 < cnt; k0 += 2) {
    int jb[2];
    double xb[2], yb[2], nb[2], mb[2], rb[2], ub[2], vb[2];
#pragma unroll
    for (int q = 0; q < 2; q++) jb[q] = (k0 + q < cnt) ? D.nbr[(k0 + q) * cap + i] : i;
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int j = jb[q];
      xb[q] = D.xnb[j]; yb[q] = D.ynb[j]; nb[q] = D.por[j]; mb[q] = D.m[j];
      rb[q] = D.rho[j]; ub[q] = D.u[j]; vb[q] = D.v[j];
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      if (k0 + q >= cnt) break;
      const Klt z = d_klt(xi - xb[q], yi - yb[q], P);
      const double gx = c1 * z.k1 + c2 * z.k2, gy = c3 * z.k1 + c4 * z.k2;
      const double nj = nb[q], mj = mb[q], rj = rb[q];
      const double du = (ub[q] / nj) - ui, dv = (vb[q] / nj) - vi;
      t1 = t1 + ((mj * du) * gx) / rj;
      t2 = t2 + ((mj * dv) * gy) / rj;
      t3 = t3 + ((0.50 * mj) * ((du * gy) + (dv * gx))) / rj;
    }
  }
  const double t4 = sqrt((2 * ((t1 * t1) + (t2 * t2))) + ((2 * t3) * (2 * t3)));
  const double dl = P.dl;
  D.nut[i] = ((0.150 * dl) * (0.150 * dl)) * t4;
  D.tke[i] = (0.080 / 1.0) * ((dl * t4) * (dl * t4));
}
__global__ void k_eddyvis(Dev D) {
  PIDX();
  eddyvis_body(D, i, n);
}

// ---------------------------------------------------------------------
// int_vel, integrator.f90:459-703 (pass 1: varts of free particles)
// ---------------------------------------------------------------------
__device__ __forceinline__ void int_vel_varts_body(const Dev& D, const int i, const int n) {
  if (i >= n || !(D.flags[i] & F_FREE)) return;
  const DevParams& P = D.P;
  const int cap = D.cap, cnt = D.ncnt[i];
  const double c1 = D.cf1[i], c2 = D.cf2[i], c3 = D.cf3[i], c4 = D.cf4[i];
  const double xi = D.xnb[i], yi = D.ynb[i];
  const double ui = D.u[i] / D.por[i], vi = D.v[i] / D.por[i];
  double t1 = 0.0, t2 = 0.0;
  // neighbours two at a time: both neighbours' fields are requested before
  // either is used (the sums stay in list order)
  for (int k0 = 0; k0 < cnt; k0 += 2) {
    int jb[2];
    double xb[2], yb[2], nb[2], mb[2], rb[2], ub[2], vb[2];
#pragma unroll
    for (int q = 0; q < 2; q++) jb[q] = (k0 + q < cnt) ? D.nbr[(k0 + q) * cap + i] : i;
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int j = jb[q];
      xb[q] = D.xnb[j]; yb[q] = D.ynb[j]; nb[q] = D.por[j]; mb[q] = D.m[j];
      rb[q] = D.rho[j]; ub[q] = D.u[j]; vb[q] = D.v[j];
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      if (k0 + q >= cnt) break;
      const Klt z = d_klt(xi - xb[q], yi - yb[q], P);
      const double gx = c1 * z.k1 + c2 * z.k2, gy = c3 * z.k1 + c4 * z.k2;
      const double nj = nb[q], vj = mb[q] / rb[q];
      t1 = t1 + (((ub[q] / nj) - ui) * gx) * vj;
      t2 = t2 + (((vb[q] / nj) - vi) * gy) * vj;
    }
  }
  const double uu = D.u[i], vv = D.v[i];
  D.varts[i] = (((1.0 * (P.dl * P.dl)) * sqrt((t1 * t1) + (t2 * t2))) * sqrt((uu * uu) + (vv * vv))) /
               ((D.sc->umax + (double)1e-6f) * D.por[i]);
}
__global__ void k_int_vel_varts(Dev D) {
  PIDX();
  int_vel_varts_body(D, i, n);
}

// pass 2: intermediate velocity (viscous + SPS + tke terms, Darcy-Forchheimer)
__device__ __forceinline__ void int_vel_body(const Dev& D, const int i, const int n) {
  if (i >= n) return;
  const DevParams& P = D.P;
  const double dt = D.sc->dt;
  const bool buf = (D.flags[i] & F_BUFFER) != 0;
  const double ni = D.por[i];
  if (D.tid[i] == TID_FLUID && !buf) {
    const double eps6 = (double)1e-6f;
    const int cap = D.cap, cnt = D.ncnt[i];
    const double c1 = D.cf1[i], c2 = D.cf2[i], c3 = D.cf3[i], c4 = D.cf4[i];
    const double xnbi = D.xnb[i], ynbi = D.ynb[i];
    const double xi = D.x[i], yi = D.y[i], ui = D.u[i], vi = D.v[i];
    const double r0i = D.rho0[i], visi = d_vis(r0i);
    const bool freei = (D.flags[i] & F_FREE) != 0;
    const double vartsi = freei ? D.varts[i] : 0.0;
    const double nuti = D.nut[i], tkei = D.tke[i];
    const double cmu_wall = visi * (1.0 / r0i);
    double us = 0.0, vs = 0.0;
    // neighbours two at a time: both neighbours' unconditional fields are
    // requested before either is used (the sums stay in list order)
    for (int k0 = 0; k0 < cnt; k0 += 2) {
    int jb[2], tb[2];
    uint8_t fb[2];
    double xnb_[2], ynb_[2], mb[2], rb[2], xb[2], yb[2], ub[2], vb[2], kb[2], r0b[2], ntb[2];
#pragma unroll
    for (int q = 0; q < 2; q++) jb[q] = (k0 + q < cnt) ? D.nbr[(k0 + q) * cap + i] : i;
    // rho0 and nu_t of the neighbour are used by fluid neighbours only (most
    // of them): read with the other fields rather than after the type test
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int j = jb[q];
      xnb_[q] = D.xnb[j]; ynb_[q] = D.ynb[j]; tb[q] = D.tid[j]; fb[q] = D.flags[j];
      mb[q] = D.m[j]; rb[q] = D.rho[j]; xb[q] = D.x[j]; yb[q] = D.y[j];
      ub[q] = D.u[j]; vb[q] = D.v[j]; kb[q] = D.tke[j]; r0b[q] = D.rho0[j]; ntb[q] = D.nut[j];
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      if (k0 + q >= cnt) break;
      const int j = jb[q];
      const Klt z = d_klt(xnbi - xnb_[q], ynbi - ynb_[q], P);
      const double gx = c1 * z.k1 + c2 * z.k2, gy = c3 * z.k1 + c4 * z.k2;
      const int tj = tb[q];
      double cmu;
      if (tj == TID_FLUID) {
        const double r0j = r0b[q], visj = d_vis(r0j);
        cmu = (((2 * visi) * visj) * ((1.0 / r0i) * (1.0 / r0j))) / ((visi / r0i) + (visj / r0j));
      } else {
        cmu = cmu_wall;
      }
      double vart;
      if (freei) vart = vartsi;
      else if ((fb[q] & F_FREE) && !buf) vart = (d_wab(z.r, P) * D.varts[j]) / P.wo;
      else vart = 0.0;
      const double pvol = mb[q] / rb[q];
      const double pd = 1.0 / ((z.r * z.r) + P.lam);
      const double dxij = xi - xb[q], dyij = yi - yb[q];
      const double uj = ub[q], vj = vb[q];
      double a1 = ((((2 * pvol) * (ui - uj)) * dxij) * gx) * pd;
      double a2 = ((((2 * pvol) * (ui - uj)) * dyij) * gy) * pd;
      us = us + (a1 + a2) * (cmu + vart);
      // the SPS viscosity of the pair (the harmonic mean is the same value
      // in both components, so it is divided once)
      double nutj = 0.0, hm = 0.0;
      bool low = false;
      if (tj > 2) {
        nutj = ntb[q];
        low = (nuti < eps6) && (nutj < eps6);
        if (!low) hm = ((2 * nuti) * nutj) / (nuti + nutj);
      }
      if (tj > 2) {
        if (low) us = us + ((a1 + a2) * (nuti + nutj)) * 0.50;
        else us = us + (a1 + a2) * hm;
      }
      a1 = ((((2 * pvol) * (vi - vj)) * dxij) * gx) * pd;
      a2 = ((((2 * pvol) * (vi - vj)) * dyij) * gy) * pd;
      vs = vs + (a1 + a2) * (cmu + vart);
      if (tj > 2) {
        if (low) vs = vs + ((a1 + a2) * (nuti + nutj)) * 0.50;
        else vs = vs + (a1 + a2) * hm;
      }
      const double dk = pvol * (kb[q] - tkei);
      double tk = dk * gx;
      us = us - ((tk * ni) * 2.0) / 3;
      tk = dk * gy;
      vs = vs - ((tk * ni) * 2.0) / 3;
    }
    }
    const double den = 1.0 - ((dt * D.resist[i]) * ni);
    us = (ui + (dt * us)) / den;
    vs = (vi + (dt * vs)) / den;
    vs = vs + ((P.g * ni) * dt);
    D.us[i] = us;
    D.vs[i] = vs;
  }
  if (buf) {
    D.us[i] = D.u[i];
    D.vs[i] = (P.g * ni) * dt;
  }
}
__global__ void __launch_bounds__(256) k_int_vel(Dev D) {
  PIDX();
  int_vel_body(D, i, n);
}

// ---------------------------------------------------------------------
// PPE assembly, isph.f90:115-456, into canonical ELL rows:
// columns ascending, the diagonal before the slot-1 value a ghost row keeps
// from the row initialisation (isph.f90:127), explicit zeros dropped.
// Raw values are written first, then divided by the diagonal (Jacobi,
// isph.f90:432-440) once the diagonal is known.
// ---------------------------------------------------------------------
__device__ __forceinline__ int ppe_assemble_row(const Dev& D, int i) {
  const DevParams& P = D.P;
  const int cap = D.cap, cnt = D.ncnt[i];
  const double dt = D.sc->dt;
  const int tidi = D.tid[i];
  const bool freei = (D.flags[i] & F_FREE) != 0;
  const double ni = D.por[i], rhoi = D.rho[i], mi = D.m[i];
  const double rho_i = rhoi * (1.0 / ni);
  const double xi = D.x[i], yi = D.y[i], usi = D.us[i], vsi = D.vs[i];
  const double xnbi = D.xnb[i], ynbi = D.ynb[i];
  const double c1 = D.cf1[i], c2 = D.cf2[i], c3 = D.cf3[i], c4 = D.cf4[i];
  const double rhoref = P.rho * P.rel_den;
  double diag = 0.0, rhs = 0.0;
  int L = 0, dpos = 0;
  bool before = true, first_written = false;
  // neighbours two at a time: both neighbours' common fields are requested
  // before either is used (entries and sums stay in list order)
  for (int k0 = 0; k0 < cnt; k0 += 2) {
  int jb[2], tb[2];
  double nb_[2], rb_[2], mb_[2], xb_[2], yb_[2], xnb_[2], ynb_[2], usb[2], vsb[2];
#pragma unroll
  for (int q = 0; q < 2; q++) jb[q] = (k0 + q < cnt) ? D.nbr[(k0 + q) * cap + i] : i;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int j = jb[q];
    tb[q] = D.tid[j]; nb_[q] = D.por[j]; rb_[q] = D.rho[j]; mb_[q] = D.m[j];
    xb_[q] = D.x[j]; yb_[q] = D.y[j]; xnb_[q] = D.xnb[j]; ynb_[q] = D.ynb[j];
    usb[q] = D.us[j]; vsb[q] = D.vs[j];
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int k = k0 + q;
    if (k >= cnt) break;
    const int j = jb[q];
    if (before && j > i) {
      // reserve the diagonal slot (and the ghost-row slot-1 value)
      before = false;
      dpos = L;
      L += 1 + ((tidi == TID_GHOST && k > 0 && !first_written) ? 1 : 0);
    }
    const int tj = tb[q];
    const double nj = nb_[q], rhoj = rb_[q];
    const double rho_j = rhoj * (1.0 / nj);
    const double pvol = mb_[q] * (1.0 / rhoj);
    const double xj = xb_[q], yj = yb_[q];
    const double ddx = xj - xi, ddy = yj - yi;
    const double dd = sqrt((ddx * ddx) + (ddy * ddy));
    const double pd = 1.0 / ((dd * dd) + P.lam);
    const Klt z = d_klt(xnbi - xnb_[q], ynbi - ynb_[q], P);
    const double dx = xi - xj, dy = yi - yj;
    bool wr = false;
    double val = 0.0, t1, t2;
    if (tidi == TID_FLUID) {
      const double lamp = (freei && tj == TID_FLUID) ? (1.0 + D.rho0[j] / rhoref) : 1.0;
      const double rho_ij = (((2 * rhoj) * rhoi) * (1.0 / (nj * ni))) *
                            (1.0 / ((rhoj * (1.0 / nj)) + (rhoi * (1.0 / ni))));
      const double gx = c1 * z.k1 + c2 * z.k2, gy = c3 * z.k1 + c4 * z.k2;
      t1 = (((2 * pvol) * gx) * dx) * pd;
      t2 = (((2 * pvol) * gy) * dy) * pd;
      val = -(t1 + t2);
      wr = true;
      diag = diag + (t1 + t2) * lamp;
      if (tj == TID_FLUID) {
        t1 = (pvol * (((usb[q] * rho_ij) / nj) - ((usi * rho_ij) / ni))) * gx;
        t2 = (pvol * (((vsb[q] * rho_ij) / nj) - ((vsi * rho_ij) / ni))) * gy;
      } else {
        t1 = (pvol * (((usb[q] * rho_i) / ni) - ((usi * rho_i) / ni))) * gx;
        t2 = (pvol * (((vsb[q] * rho_i) / ni) - ((vsi * rho_i) / ni))) * gy;
      }
      rhs = rhs + (t1 + t2) / dt;
    } else if (tidi == TID_GHOST) {
      if (tj <= 2) {
        t1 = (((2 * pvol) * z.k1) * dx) * pd;
        t2 = (((2 * pvol) * z.k2) * dy) * pd;
        val = -(t1 + t2);
        wr = true;
        diag = diag + (t1 + t2);
      }
    } else if (tidi == TID_FIXED) {
      if (tj == TID_FLUID) {
        const double gxj = D.cf1[j] * z.k1 + D.cf2[j] * z.k2;
        const double gyj = D.cf3[j] * z.k1 + D.cf4[j] * z.k2;
        t1 = ((((2 * mi) * gxj) * dx) * pd) / rhoi;
        t2 = ((((2 * mi) * gyj) * dy) * pd) / rhoi;
        val = -(t1 + t2);
        wr = true;
        diag = diag + (t1 + t2);
        t1 = ((mi * (((D.us[j] * rho_j) / nj) - ((usi * rho_j) / ni))) * gxj) / rhoi;
        t2 = ((mi * (((D.vs[j] * rho_j) / nj) - ((vsi * rho_j) / ni))) * gyj) / rhoi;
        rhs = rhs + (t1 + t2) / dt;
      } else if (tj <= 2) {
        t1 = (((2 * pvol) * z.k1) * dx) * pd;
        t2 = (((2 * pvol) * z.k2) * dy) * pd;
        val = -(t1 + t2);
        wr = true;
        diag = diag + (t1 + t2);
      }
    }
    if (k == 0) first_written = wr;
    if (wr) {
      D.acol[L * cap + i] = j;
      D.aval[L * cap + i] = val;
      L++;
    }
  }
  }
  if (before) {
    dpos = L;
    L += 1 + ((tidi == TID_GHOST && cnt > 0 && !first_written) ? 1 : 0);
  }
  // A particle with no neighbours gets the identity row, so p = 0 (its rhs
  // is 0): the diagonal value 1 the reference leaves commented out at
  // isph.f90:425-427, where its live code divides 0 by 0 (isph.f90:432-440).
  if (cnt == 0) diag = 1.0;
  const double t1 = diag;
  if (!(isfinite(t1) && t1 != 0.0)) atomicOr(&D.sc->err, ERR_DIAG);
  D.acol[dpos * cap + i] = i;
  D.aval[dpos * cap + i] = diag / t1;
  const bool quirk = (tidi == TID_GHOST && cnt > 0 && D.nbr[i] < i && !first_written);
  if (quirk) {
    D.acol[(dpos + 1) * cap + i] = i;
    D.aval[(dpos + 1) * cap + i] = 1.0 / t1;
  }
  for (int k = 0; k < L; k++) {
    if (k == dpos || (quirk && k == dpos + 1)) continue;
    D.aval[k * cap + i] = D.aval[k * cap + i] / t1;
  }
  D.alen[i] = L;
  D.b[i] = rhs / t1;
  return L;
}

// the stored entries are counted per workgroup: one atomic per workgroup
// rather than one per row on a single address
__device__ __forceinline__ void ppe_assemble_body(const Dev& D, const int i, const int n) {
  const int L = (i < n) ? ppe_assemble_row(D, i) : 0;
  const unsigned long long s = block_sum_u64((unsigned long long)L);
  if (threadIdx.x == 0 && s != 0) atomicAdd(&D.sc->nnz, s);
}
__global__ void __launch_bounds__(256) k_ppe_assemble(Dev D) {
  PIDX();
  ppe_assemble_body(D, i, n);
}

// ---------------------------------------------------------------------
// comp_vel, integrator.f90:705-843 (with the running maxima of timestep)
// ---------------------------------------------------------------------
// bits of |v| for the running maxima (0 for NaN, which they leave out)
__device__ __forceinline__ unsigned long long max_bits(double v) {
  const double a = fabs(v);
  return (a == a) ? (unsigned long long)__double_as_longlong(a) : 0ull;
}

__device__ __forceinline__ void comp_vel_row(const Dev& D, int i) {
  const DevParams& P = D.P;
  const double dt = D.sc->dt;
  const bool buf = (D.flags[i] & F_BUFFER) != 0;
  if (D.tid[i] == TID_FLUID && !buf) {
    const int cap = D.cap, cnt = D.ncnt[i];
    const double c1 = D.cf1[i], c2 = D.cf2[i], c3 = D.cf3[i], c4 = D.cf4[i];
    const double xnbi = D.xnb[i], ynbi = D.ynb[i];
    const double ni = D.por[i], rhoi = D.rho[i], pi = D.p[i];
    double u = 0.0, v = 0.0;
    // neighbours two at a time: both neighbours' fields are requested before
    // either is used (the sums stay in list order)
    for (int k0 = 0; k0 < cnt; k0 += 2) {
      int jb[2];
      double xb[2], yb[2], nb[2], rb[2], pb[2], mb[2];
      int tb[2];
#pragma unroll
      for (int q = 0; q < 2; q++) jb[q] = (k0 + q < cnt) ? D.nbr[(k0 + q) * cap + i] : i;
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const int j = jb[q];
        xb[q] = D.xnb[j]; yb[q] = D.ynb[j]; nb[q] = D.por[j]; rb[q] = D.rho[j];
        pb[q] = D.p[j]; mb[q] = D.m[j]; tb[q] = D.tid[j];
      }
#pragma unroll
      for (int q = 0; q < 2; q++) {
      if (k0 + q >= cnt) break;
      const Klt z = d_klt(xnbi - xb[q], ynbi - yb[q], P);
      const double gx = c1 * z.k1 + c2 * z.k2, gy = c3 * z.k1 + c4 * z.k2;
      const double nj = nb[q], rhoj = rb[q], pj = pb[q];
      const double V = mb[q] / rhoj;
      double t1;
      if (tb[q] == TID_FLUID) {
        const double a = ((pj * rhoi) / ni) + ((pi * rhoj) / nj);
        const double den = rhoi * ((rhoi / ni) + (rhoj / nj));
        t1 = (2 * (((V * (ni * ni)) * a) * gx)) / den;
        u = u - t1 * dt;
        t1 = (2 * (((V * (ni * ni)) * a) * gy)) / den;
        v = v - t1 * dt;
      } else {
        t1 = (((V * (ni * ni)) * (pj + pi)) * gx) / rhoi;
        u = u - t1 * dt;
        t1 = (((V * (ni * ni)) * (pj + pi)) * gy) / rhoi;
        v = v - t1 * dt;
      }
      }
    }
    u = D.us[i] + u;
    v = D.vs[i] + v;
    D.u[i] = u;
    D.v[i] = v;
    D.vel[i] = sqrt((u * u) + (v * v)) / ni;
  }
  if (buf) {
    D.u[i] = D.us[i];
    D.v[i] = 0.0;
  }
}

__device__ __forceinline__ void comp_vel_body(const Dev& D, const int i, const int n) {
  if (i < n) comp_vel_row(D, i);
}
__global__ void __launch_bounds__(256) k_comp_vel(Dev D) {
  PIDX();
  comp_vel_body(D, i, n);
}

// The per-cell maxima of comp_vel (integrator.f90:831-841): a cell that
// holds particles takes the maxima of |vel| and |nu_t| over them; an empty
// cell keeps the values of the last comp_vel that found particles in it.
// vel is that of the non-buffer fluid particles just updated (walls
// contribute 0, as their never-written vel does in the reference; so does a
// buffer particle, whose slot the reference reads stale).  With init the
// velocities count as 0 and empty cells are cleared: the state before the
// first step.  Each block of 256 cells leaves the maxima over its cells for
// timestep (one thread per cell; non-negative doubles compare as their bits).
__device__ __forceinline__ void cell_max_cell(const Dev& D, int k, int ncell, int init,
                                              int blk) {
  unsigned long long vb = 0ull, nb = 0ull;
  if (k < ncell) {
    const int i0 = D.cstart[k], i1 = D.cstart[k + 1];
    if (i1 > i0) {
      // eight rows' fields requested before any is used (a maximum of
      // non-negative bit patterns does not depend on the order)
      for (int b0 = i0; b0 < i1; b0 += 8) {
        uint8_t tb[8], fb[8];
        double vl[8], nt[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int i = (b0 + u < i1) ? b0 + u : i0;
          tb[u] = D.tid[i];
          fb[u] = D.flags[i];
          vl[u] = D.vel[i];
          nt[u] = D.nut[i];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          if (b0 + u >= i1) break;
          if (!init && tb[u] == TID_FLUID && !(fb[u] & F_BUFFER)) {
            const unsigned long long b = max_bits(vl[u]);
            vb = (b > vb) ? b : vb;
          }
          const unsigned long long b = max_bits(nt[u]);
          nb = (b > nb) ? b : nb;
        }
      }
      D.cmaxv[k] = vb;
      D.cmaxn[k] = nb;
    } else if (init) {
      D.cmaxv[k] = 0ull;
      D.cmaxn[k] = 0ull;
    } else {
      vb = D.cmaxv[k];
      nb = D.cmaxn[k];
    }
  }
  block_max2_u64(vb, nb);
  if (threadIdx.x == 0) {
    D.maxb[blk] = vb;
    D.maxb[D.nmaxb + blk] = nb;
  }
}

__global__ void __launch_bounds__(256) k_cell_max(Dev D, int ncell, int init) {
  cell_max_cell(D, blockIdx.x * blockDim.x + threadIdx.x, ncell, init, blockIdx.x);
}

// comp_pos, integrator.f90:845-883
__device__ __forceinline__ void comp_pos_body(const Dev& D, const int i, const int n) {
  if (i >= n || D.tid[i] != TID_FLUID) return;
  const double dt = D.sc->dt;
  const double x = ((D.x[i] + D.xs[i]) + ((dt * D.u[i]) / D.por[i])) / 2.0;
  const double y = ((D.y[i] + D.ys[i]) + ((dt * D.v[i]) / D.por[i])) / 2.0;
  D.x[i] = x;
  D.y[i] = y;
  if ((D.flags[i] & F_BUFFER) && (x > D.P.buf_xmin) && (x < D.P.buf_xmax))
    D.flags[i] = D.flags[i] & (uint8_t)~F_BUFFER;
}
__global__ void k_comp_pos(Dev D) {
  PIDX();
  comp_pos_body(D, i, n);
}

// freesurf, isph.f90:49-86: surface particles mark their neighbours
// (plain byte stores of the same value; the order of the stores is irrelevant)
__device__ __forceinline__ void freesurf_body(const Dev& D, const int i, const int n) {
  if (i >= n) return;
  if (D.tid[i] == TID_FLUID && !(D.flags[i] & F_BUFFER) && D.gradvx[i] <= D.P.fs_thresh) {
    const int cap = D.cap, cnt = D.ncnt[i];
    for (int k = 0; k < cnt; k++) D.vic[D.nbr[k * cap + i]] = 1;
  }
}
__global__ void k_freesurf(Dev D) {
  PIDX();
  freesurf_body(D, i, n);
}

// ---------------------------------------------------------------------
// opt2_shift, part_shift.f90:12-191.  The surface normal and pint are zero
// at the point of use (part_shift.f90:67-69), so the shift sums reduce to
// -csh*h^2*frac*gradW*V; the stored shift is the capped one, the velocity
// correction uses the uncapped shift.
// ---------------------------------------------------------------------
__device__ __forceinline__ void opt2_shift_body(const Dev& D, const int i, const int n) {
  if (i >= n) return;
  if (!(D.tid[i] == TID_FLUID && !(D.flags[i] & F_BUFFER))) return;
  const int cnt = D.ncnt[i];
  if (cnt == 0) {
    // no neighbour: no shift.  The reference zeroes its local shifts
    // (part_shift.f90:32-35) but stores them only for count > 0
    // (part_shift.f90:38-161), so its apply step (part_shift.f90:179-182)
    // would add the slot's saved position and intermediate velocity.
    D.xs[i] = 0.0;
    D.ys[i] = 0.0;
    D.us[i] = 0.0;
    D.vs[i] = 0.0;
    return;
  }
  const DevParams& P = D.P;
  const int cap = D.cap;
  const double c1 = D.cf1[i], c2 = D.cf2[i], c3 = D.cf3[i], c4 = D.cf4[i];
  const double xnbi = D.xnb[i], ynbi = D.ynb[i];
  const double frac = (D.flags[i] & F_FREE) ? 0.01 : 1.0;
  const double coef = (P.csh * (P.h * P.h)) * frac;
  const double ui = D.u[i], vi = D.v[i];
  double xsa = 0.0, ysa = 0.0, a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0;
  // neighbours two at a time: both neighbours' fields are requested before
  // either is used (the sums stay in list order)
  for (int k0 = 0; k0 < cnt; k0 += 2) {
    int jb[2];
    double xb[2], yb[2], mb[2], rb[2], ub[2], vb[2];
#pragma unroll
    for (int q = 0; q < 2; q++) jb[q] = (k0 + q < cnt) ? D.nbr[(k0 + q) * cap + i] : i;
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int j = jb[q];
      xb[q] = D.xnb[j]; yb[q] = D.ynb[j]; mb[q] = D.m[j]; rb[q] = D.rho[j];
      ub[q] = D.u[j]; vb[q] = D.v[j];
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      if (k0 + q >= cnt) break;
      const Klt z = d_klt(xnbi - xb[q], ynbi - yb[q], P);
      const double V = mb[q] / rb[q];
      xsa = xsa - ((coef * z.k1) * V);
      ysa = ysa - ((coef * z.k2) * V);
      const double gx = c1 * z.k1 + c2 * z.k2, gy = c3 * z.k1 + c4 * z.k2;
      const double du = ub[q] - ui, dv = vb[q] - vi;
      a1 = a1 + (gx * du) * V;
      a2 = a2 + (gy * du) * V;
      b1 = b1 + (gx * dv) * V;
      b2 = b2 + (gy * dv) * V;
    }
  }
  const double capv = P.maxshift * P.dl;
  const double ax = fabs(xsa), ay = fabs(ysa);
  D.xs[i] = (xsa == 0.0) ? 0.0 : (((capv < ax) ? capv : ax) * xsa) / ax;
  D.ys[i] = (ysa == 0.0) ? 0.0 : (((capv < ay) ? capv : ay) * ysa) / ay;
  D.us[i] = a1 * xsa + a2 * ysa;
  D.vs[i] = b1 * xsa + b2 * ysa;
}
__global__ void __launch_bounds__(256) k_opt2_shift(Dev D) {
  PIDX();
  opt2_shift_body(D, i, n);
}

__device__ __forceinline__ void shift_apply_body(const Dev& D, const int i, const int n) {
  if (i >= n) return;
  if (D.tid[i] == TID_FLUID && !(D.flags[i] & F_BUFFER)) {
    D.x[i] = D.x[i] + D.xs[i];
    D.y[i] = D.y[i] + D.ys[i];
    D.u[i] = D.u[i] + D.us[i];
    D.v[i] = D.v[i] + D.vs[i];
  }
}
__global__ void k_shift_apply(Dev D) {
  PIDX();
  shift_apply_body(D, i, n);
}

// densityupdate, scalar.f90:399-426
__device__ __forceinline__ void densityupdate_body(const Dev& D, const int i, const int n) {
  if (i >= n || D.tid[i] != TID_FLUID) return;
  const double rhomax = D.P.rho * D.P.rel_den, rhomin = D.P.rhomin;
  const double r0 = (0.5 * (rhomax + rhomin)) + (D.c[i] * (rhomax - rhomin));
  D.rho0[i] = r0;
  D.m[i] = r0 * D.vol0[i];
  D.rho[i] = r0 * D.por[i];
}
__global__ void k_densityupdate(Dev D) {
  PIDX();
  densityupdate_body(D, i, n);
}

// timestep, integrator.f90:885-953 (one workgroup: the maxima over every
// cell, from the cell blocks' maxima, then one thread)
__device__ __forceinline__ double dmin2(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ void timestep_block(const Dev& D) {
  unsigned long long mv = 0, mn = 0;
  for (int b0 = threadIdx.x; b0 < D.nmaxb; b0 += 4 * blockDim.x) {
    unsigned long long qv[4], qn[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int b = b0 + u * (int)blockDim.x;
      qv[u] = (b < D.nmaxb) ? D.maxb[b] : 0ull;
      qn[u] = (b < D.nmaxb) ? D.maxb[D.nmaxb + b] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      mv = (qv[u] > mv) ? qv[u] : mv;
      mn = (qn[u] > mn) ? qn[u] : mn;
    }
  }
  block_max2_u64(mv, mn);
  if (threadIdx.x != 0) return;
  const DevParams& P = D.P;
  DevScalars* S = D.sc;
  S->umax = __longlong_as_double((long long)mv);
  S->numax = __longlong_as_double((long long)mn);
  const double dl = P.dl, sig1 = P.sig1;
  if (!P.fixed_dt) {
    const double t = S->t;
    double dtn;
    if (t <= 1.0) {
      dtn = dmin2(dmin2(sig1 * dl / S->umax, (sig1 * (dl * dl)) / ((P.mu / P.rho) + S->numax)), 0.0010);
    } else if ((t > 1.0) && (t <= 2.50)) {
      float a = (float)(sig1 * dl / S->umax), b = (float)(((sig1 * (dl * dl)) * P.rho) / P.mu);
      float mn = a;
      if (b < mn) mn = b;
      if (0.002f < mn) mn = 0.002f;
      dtn = (double)mn;
    } else if ((t > 3.0) && (t <= 4.0)) {
      float a = (float)(sig1 * dl / S->umax), b = (float)(((sig1 * (dl * dl)) * P.rho) / P.mu);
      float mn = a;
      if (b < mn) mn = b;
      if (0.004f < mn) mn = 0.004f;
      dtn = (double)mn;
    } else {
      dtn = dmin2(dmin2(sig1 * dl / S->umax, (sig1 * (dl * dl)) / ((P.mu / P.rho) + S->numax)), 0.0050);
    }
    S->dt = dtn;
  }
  if (!isfinite(S->numax) || isnan(S->numax)) S->numax = 0.0;
  double dtsol = (sig1 * (dl * dl)) / ((double)1e-9f + S->numax / P.tschmidt);
  int solsteps = 1;
  if (dtsol < S->dt) {
    solsteps = (int)ceil(S->dt / dtsol);
    dtsol = S->dt / solsteps;
  }
  S->dtsol = (dtsol < S->dt) ? dtsol : S->dt;
  S->solsteps = solsteps;
}
__global__ void __launch_bounds__(256) k_timestep(Dev D) {
  if (blockIdx.x == 0) timestep_block(D);
}

// ---------------------------------------------------------------------
// scalart_anisotropic / scalarupdate, scalar.f90:242-397.  The neighbour
// list (and its r) is from the last neighbour search; positions are current.
// ---------------------------------------------------------------------
__device__ __forceinline__ void scalart_body(const Dev& D, const int i, const int n) {
  if (i >= n || D.tid[i] != TID_FLUID) return;
  const DevParams& P = D.P;
  const double al = 0.0025, at = 0.00025, Dm = (double)1e-9f, thr = 0.7;
  const double eps6 = (double)1e-6f;
  const double dt = D.sc->dt;
  const double capv = (0.12 * (P.dl * P.dl)) / dt;
  const int cap = D.cap, cnt = D.ncnt[i];
  const double xnbi = D.xnb[i], ynbi = D.ynb[i], xi = D.x[i], yi = D.y[i];
  const double ui = D.u[i], vi_ = D.v[i], ni = D.por[i], ci1 = D.c[i];
  const double nuti = D.nut[i], vol0i = D.vol0[i];
  const double vi = sqrt((ui * ui) + (vi_ * vi_)) / ni;
  double cd = 0.0;
  // neighbours two at a time: both neighbours' fields are requested before
  // either is used (the sum stays in list order)
  for (int k0 = 0; k0 < cnt; k0 += 2) {
  int tb[2];
  double xnb_[2], ynb_[2], ub[2], vb[2], nb[2], cb[2], xb[2], yb[2], nutb[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int j = (k0 + q < cnt) ? D.nbr[(k0 + q) * cap + i] : i;
    tb[q] = D.tid[j]; xnb_[q] = D.xnb[j]; ynb_[q] = D.ynb[j]; ub[q] = D.u[j]; vb[q] = D.v[j];
    nb[q] = D.por[j]; cb[q] = D.c[j]; xb[q] = D.x[j]; yb[q] = D.y[j]; nutb[q] = D.nut[j];
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    if (k0 + q >= cnt) break;
    if (tb[q] != TID_FLUID) continue;
    const double ex = xnbi - xnb_[q], ey = ynbi - ynb_[q];
    const double r = sqrt((ex * ex) + (ey * ey));
    const double uj = ub[q], vj_ = vb[q], nj = nb[q], c2 = cb[q];
    const double dx = xi - xb[q], dy = yi - yb[q];
    double t1x;
    if ((ni < thr) && (nj < thr)) {
      const double vj = sqrt((uj * uj) + (vj_ * vj_)) / nj;
      const double Dxxi = ui / ((sqrt((al * vi) + Dm) * vi) * ni);
      const double Dxyi = vi_ / ((sqrt((al * vi) + Dm) * vi) * ni);
      const double Dyxi = (-vi_) / ((sqrt((at * vi) + Dm) * vi) * ni);
      const double Dyyi = ui / ((sqrt((at * vi) + Dm) * vi) * ni);
      const double Dxxj = uj / ((sqrt((al * vj) + Dm) * vj) * nj);
      const double Dxyj = vj_ / ((sqrt((al * vj) + Dm) * vj) * nj);
      const double Dyxj = (-vj_) / ((sqrt((at * vj) + Dm) * vj) * nj);
      const double Dyyj = uj / ((sqrt((at * vj) + Dm) * vj) * nj);
      const double pd = 1.0 / r;
      const double A = ((((Dxxi + Dxxj) * dx) * pd) * 0.5) + ((((Dxyi + Dxyj) * dy) * pd) * 0.5);
      const double B = ((((Dyxi + Dyxj) * dx) * pd) * 0.5) + ((((Dyyi + Dyyj) * dy) * pd) * 0.5);
      t1x = (A * A) + (B * B);
      t1x = 1.0 / t1x;
      t1x = (t1x <= capv) ? t1x : capv;
    } else {
      const double nutj = nutb[q];
      const double Dxxi = Dm + nuti / P.tschmidt;
      const double Dxxj = Dm + nutj / P.tschmidt;
      if ((nuti < eps6) && (nutj < eps6)) t1x = (Dxxi + Dxxj) * 0.50;
      else t1x = ((2 * Dxxi) * Dxxj) / (Dxxi + Dxxj);
    }
    const double t3 = (vol0i * (ni + nj)) / (ni * nj);
    double hwx, hwy;
    d_wabd2(dx, dy, r, P, hwx, hwy);
    const double den = (r * r) + P.lam;
    cd = (cd + (((((((t1x * 1.0) * t3) * (ci1 - c2)) * dx) * hwx) / den))) +
         (((((((t1x * 1.0) * t3) * (ci1 - c2)) * dy) * hwy) / den));
  }
  }
  D.cdiff[i] = cd;
}
__global__ void __launch_bounds__(256) k_scalart(Dev D) {
  PIDX();
  scalart_body(D, i, n);
}

// stime < 0: use the sub-step length dtsol from the last timestep
__device__ __forceinline__ void scalarupdate_body(const Dev& D, const int i, const int n, double stime) {
  if (i >= n || D.tid[i] != TID_FLUID) return;
  const double st = (stime < 0.0) ? D.sc->dtsol : stime;
  D.c[i] = D.c[i] + D.cdiff[i] * st;
  if ((D.flags[i] & F_BUFFER) && (D.x[i] < 1.6)) D.c[i] = 0.5;
}
__global__ void k_scalarupdate(Dev D, double stime) {
  PIDX();
  scalarupdate_body(D, i, n, stime);
}
