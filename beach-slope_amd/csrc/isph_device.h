// isph_device.h -- device data layout and the per-particle math of the ISPH
// pipeline for gfx950.
//
// Layout in HBM (one context = one scenario):
//   * particle state as structure-of-arrays in cell-sorted order; a particle's
//     index is its PPE row (the reference's matid - 1 after resetid,
//     isph.f90:88-113);
//   * neighbour lists and PPE rows in ELL form, slot-major
//     (a[k*cap + i] for slot k of particle i), so that a wavefront reading
//     slot k of 64 consecutive particles touches one contiguous run;
//   * cell starts for the cell-linked list, (cell, transfer-class) bins for
//     the rebin.
// Every expression below is evaluated in the order it is written in the
// reference Fortran (files cited per function); the library is compiled with
// -ffp-contract=off so no FMA is formed and every rounding happens where the
// Fortran rounds.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NBR_CAP 64
#define ROW_CAP (NBR_CAP + 2)
#define RED_BLOCK 256
#define NCLS 10     // transfer classes per destination cell
#define CLS_DROP 15

struct DevParams {
  int cellx, celly;
  double h, dl, lam, rho, rel_den, rhomin, g, mu, por, d50, lamk, csh, maxshift;
  double fs_thresh, sig1, tschmidt, xlcutoff, ytcutoff, buf_xmin, buf_xmax;
  double tol, inner_tol;
  int fixed_dt, max_iter, max_outer;
  // constants derived on the host with the same expressions as the reference
  double cs;        // incr*h1 = 4h (setup.f90:120-123)
  double k1, k2;    // 10/(28*3.14 h^2), 1/(3*3.14 h^2), real4 products (kernel.f90:61-62)
  double two_h, lim2, wo;
  // weakly compressible variant: sound speed, Tait exponent, delta-SPH
  // coefficient, still-water level of the hydrostatic density
  double c0, gam, delt, swl;
};

// Device-resident scalars (the module variables t, dt, dtsol, solsteps,
// umax, numax of initialize.f90:21-60) plus solver state.
struct DevScalars {
  double t, dt, dtsol, umax, numax;
  unsigned long long umax_bits, numax_bits;  // unused (layout kept)
  int solsteps, n;
  int stop_seq;  // precision form: the iteration (launch sequence number) that stopped
  int err;
  // BiCGStab state
  double rho, alpha, omega, beta, rr, bnorm, stop, rel;
  int iters, done, conv, brk;
  int fin, restarts;  // precision form: scenario finished; restarts from the FP64 residual
  int far;            // precision form: a column offset of the rows exceeds 16 bits
  unsigned int counter[4];
  unsigned long long nnz;  // stored PPE entries of the last assembly
};

// error bits in DevScalars::err
#define ERR_NBR_CAP 1
#define ERR_DIAG 2

struct Dev {
  DevParams P;
  DevScalars* sc;
  int cap;
  int nblk;  // row blocks of RED_BLOCK rows covering the host's upper bound of n
  // per cell: the maxima of |velocity| and |nu_t| over the cell's particles
  // at the last comp_vel in which the cell held particles (the reference's
  // dpcell%maxvel / maxeddy, integrator.f90:831-841), as bits of
  // non-negative doubles; an emptied cell keeps its values
  unsigned long long *cmaxv, *cmaxn;
  // per block of 256 cells: the maxima over its cells, at [b] and [nmaxb + b],
  // folded by timestep (integrator.f90:914-924)
  unsigned long long* maxb;
  int nmaxb;
  // gathered particle state (ping-pong buffers swapped by the host)
  double *x, *y, *u, *v, *us, *vs, *xs, *ys, *p, *m, *rho, *rho0, *vol0, *c;
  uint8_t *tid, *flags, *vic;
  int32_t* cell;
  double *x2, *y2, *u2, *v2, *us2, *vs2, *xs2, *ys2, *p2, *m2, *rho2, *rho02, *vol02, *c2;
  uint8_t *tid2, *flags2, *vic2;
  int32_t* cell2;
  // derived properties
  double *por, *resist, *gradvx, *cf1, *cf2, *cf3, *cf4, *nut, *tke, *varts, *cdiff, *vel;
  double *xnb, *ynb;
  double* tden;  // continuity update of the weakly compressible variant
  int32_t *ncnt, *nbr;
  int32_t *cstart, *bincnt, *binstart, *pbin, *prank, *scan_tmp;
  // solids
  double *sx, *sy, *sm, *srho;
  int32_t* sstart;
  // PPE rows and solver vectors
  int32_t *alen, *acol;
  double *aval, *b, *sol;
  uint32_t* av16;   // precision form: FP16 values of the V-cycle's level-0 matrix,
                    // slots 2w and 2w+1 of row i in word [w * cap + i] (from setup)
  uint32_t* aoff;   // precision form: column offsets j - i as 16-bit pairs,
                    // slots 2w and 2w+1 of row i in word [w * cap + i]
  double *r, *rh, *pp, *t1, *sv, *t2, *part;
  float *r32, *rh32, *pp32, *sv32, *d32;
};

// A read through the global address space.  A pointer loaded from a device
// table is generic, and a read through it is a flat load, which waits on both
// the vector and the scalar counters; through the global space it is a scalar
// load when uniform and no store precedes it, else a vector load.
template <typename T>
__device__ __forceinline__ T ld_global(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

// ---------------------------------------------------------------------
// Row-block mapping of the solver's row kernels.  Workgroups are dispatched
// to the 8 XCDs round robin (workgroup b runs on XCD b mod 8) and each XCD
// has its own L2.  The grid is launched as 8 * per workgroups, and workgroup
// b takes row block (b mod 8) * per + b / 8, so every XCD sweeps one
// contiguous eighth of the rows: the neighbour gathers of a row then hit the
// L2 of the XCD that loaded the neighbouring rows.  Workgroups whose block is
// at or beyond nblk return at once (all their threads together).
// ---------------------------------------------------------------------
__device__ __forceinline__ int row_block() {
  const int per = gridDim.x >> 3;
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

// Sum over the ELL row i (L entries, slot k at k*cap + i) of val * gather(col),
// U entries at a time: the column and value loads of U entries are issued
// together, then their U gathers, then the products are added in entry
// order.  The result is the sequential sum acc = acc + val_k * x_k, rounded
// exactly as the one-entry-at-a-time loop.
template <int U, typename A, typename VAL, typename G>
__device__ __forceinline__ A ell_row_sum(const int32_t* __restrict__ col,
                                         const VAL* __restrict__ val, int cap, int i, int L,
                                         G gather) {
  A acc = (A)0;
  for (int k0 = 0; k0 < L; k0 += U) {
    int c[U];
    VAL v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool in = k0 + u < L;
      c[u] = in ? col[(k0 + u) * cap + i] : i;
      v[u] = in ? val[(k0 + u) * cap + i] : (VAL)0;
    }
    A g[U];
#pragma unroll
    for (int u = 0; u < U; u++) g[u] = (k0 + u < L) ? gather(c[u]) : (A)0;
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k0 + u < L) acc = acc + v[u] * g[u];
  }
  return acc;
}
#define ELL_U 8

// The same walk over 16-bit column offsets j - i packed two to a word (a
// quarter of the index bytes and half the index loads of the int32
// columns); used when every offset of the rows fits 16 bits.  U is even.
template <int U, typename A, typename VAL, typename G>
__device__ __forceinline__ A ell_row_sum_off(const uint32_t* __restrict__ off,
                                             const VAL* __restrict__ val, int cap, int i, int L,
                                             G gather) {
  A acc = (A)0;
  for (int k0 = 0; k0 < L; k0 += U) {
    int c[U];
    VAL v[U];
#pragma unroll
    for (int q = 0; q < U / 2; q++) {
      const uint32_t wd = (k0 + 2 * q < L) ? off[((k0 >> 1) + q) * cap + i] : 0u;
      c[2 * q] = i + (int)(int16_t)(wd & 0xffffu);
      c[2 * q + 1] = i + (int)(int16_t)(wd >> 16);
    }
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = (k0 + u < L) ? val[(k0 + u) * cap + i] : (VAL)0;
    A g[U];
#pragma unroll
    for (int u = 0; u < U; u++) g[u] = (k0 + u < L) ? gather(c[u]) : (A)0;
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k0 + u < L) acc = acc + v[u] * g[u];
  }
  return acc;
}

// Row sum for the V-cycle's level-0 matrix: FP16 values two to a word, and
// the packed 16-bit column offsets (OFF) or the int32 columns; the products
// are summed in FP32 in entry order.  U is even.
template <bool OFF, int U, typename G>
__device__ __forceinline__ float ell_row_sum_h(const Dev& D, int i, G gather) {
  const int cap = D.cap, L = D.alen[i];
  float acc = 0.0f;
  for (int k0 = 0; k0 < L; k0 += U) {
    int c[U];
    float v[U];
#pragma unroll
    for (int q = 0; q < U / 2; q++) {
      const bool in = k0 + 2 * q < L;
      const uint32_t hv = in ? D.av16[((k0 >> 1) + q) * cap + i] : 0u;
      v[2 * q] = __half2float(__ushort_as_half((unsigned short)(hv & 0xffffu)));
      v[2 * q + 1] = __half2float(__ushort_as_half((unsigned short)(hv >> 16)));
      if (OFF) {
        const uint32_t wd = in ? D.aoff[((k0 >> 1) + q) * cap + i] : 0u;
        c[2 * q] = i + (int)(int16_t)(wd & 0xffffu);
        c[2 * q + 1] = i + (int)(int16_t)(wd >> 16);
      } else {
        c[2 * q] = in ? D.acol[(k0 + 2 * q) * cap + i] : i;
        c[2 * q + 1] = (k0 + 2 * q + 1 < L) ? D.acol[(k0 + 2 * q + 1) * cap + i] : i;
      }
    }
    float g[U];
#pragma unroll
    for (int u = 0; u < U; u++) g[u] = (k0 + u < L) ? gather(c[u]) : 0.0f;
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k0 + u < L) acc = acc + v[u] * g[u];
  }
  return acc;
}

// row sum over int32 columns or packed offsets
#ifndef ELL_UP
#define ELL_UP 8   // entries per batch of the precision form's FP32 row walks
#endif
template <bool OFF, int U, typename A, typename VAL, typename G>
__device__ __forceinline__ A ell_row_sum_any(const Dev& D, const VAL* __restrict__ val, int i,
                                             G gather) {
  if (OFF) return ell_row_sum_off<U, A>(D.aoff, val, D.cap, i, D.alen[i], gather);
  return ell_row_sum<U, A>(D.acol, val, D.cap, i, D.alen[i], gather);
}

// ---------------------------------------------------------------------
// kernel.f90 restatement (device)
// ---------------------------------------------------------------------
// Wab, kernel.f90:34-66 (hWab, kernel.f90:164-184, is identical).
__device__ __forceinline__ double d_wab(double r, const DevParams& P) {
  const double h = P.h;
  double t1, t2;
  if (r < h) {
    t1 = (4.0 - ((6.0 * (r * r)) / (h * h))) + ((3.0 * ((r * r) * r)) / ((h * h) * h));
    const double q = r / h;
    t2 = (((q * q) * q) - ((6.0 * r) / h)) + 6.0;
  } else {
    const double a = 2.0 - r / h;
    t1 = (a * a) * a;
    t2 = t1;
  }
  return ((0.1 * t1) * P.k1) + (((1.0 - 0.1) * t2) * P.k2);
}

// Wabx and Waby (kernel.f90:94-162; hWabx / hWaby, kernel.f90:204-248, are
// identical) of one pair (p1 = j, p2 = i), with (dx, dy) = coordinates of i
// minus those of j.  Each component is the reference's expression
//   r < h:  t1 = ((9 r / h^3) - 12 / h^2) d,  t2 = ((3 r / h^3) - 6 / (r h)) d
//   else:   t1 = t2 = ((3 / (r h)) (2 - r / h)^2) (-d)
//   W'     = (0.1 t1) k1 + (0.9 t2) k2,
// rounded as written; the bracketed factors depend on r alone, so they are
// formed once (their divisions are most of the pair's FP64 work) and
// multiply both components, which gives each component the same bits.  The
// lanes of a wavefront take both branches, so the divisions of the two
// branches share instructions: lane by lane the first divides 9r by h^3 or
// r by h, the second 3r by h^3 or 3 by r h, each its own branch's operands.
__device__ __forceinline__ void d_wabd2(double dx, double dy, double r, const DevParams& P,
                                        double& kx, double& ky) {
  const double h = P.h;
  const bool nr = r < h;
  const double hhh = (h * h) * h, rh = r * h;
  const double qa = (nr ? 9.0 * r : r) / (nr ? hhh : h);      // 9r / h^3  |  r / h
  const double qb = (nr ? 3.0 * r : 3.0) / (nr ? hhh : rh);   // 3r / h^3  |  3 / (r h)
  double t1x, t2x, t1y, t2y;
  if (nr) {
    const double f1 = qa - (12.0 / (h * h));
    const double f2 = qb - (6.0 / rh);
    t1x = f1 * dx;
    t2x = f2 * dx;
    t1y = f1 * dy;
    t2y = f2 * dy;
  } else {
    const double a = 2.0 - qa;
    const double f = qb * (a * a);
    t1x = f * (-dx);
    t2x = t1x;
    t1y = f * (-dy);
    t2y = t1y;
  }
  kx = ((0.1 * t1x) * P.k1) + (((1.0 - 0.1) * t2x) * P.k2);
  ky = ((0.1 * t1y) * P.k1) + (((1.0 - 0.1) * t2y) * P.k2);
}

// kernel value and gradient pair for one neighbour (both components share r)
struct Klt {
  double r, k1, k2;
};
__device__ __forceinline__ Klt d_klt(double dx, double dy, const DevParams& P) {
  Klt o;
  o.r = sqrt((dx * dx) + (dy * dy));  // dist, functions.f90:28-42
  d_wabd2(dx, dy, o.r, P, o.k1, o.k2);
  return o;
}

// vis, porous.f90:12-30 / integrator.f90:12-30
__device__ __forceinline__ double d_vis(double oden) {
  const double sp = oden - 1000.0;
  return ((1.002 + (0.001652 * sp)) + (0.0000083 * (sp * sp))) * 0.001;
}

__device__ __forceinline__ void cell_bounds(int key, const DevParams& P, int& ci, int& cj,
                                            double& xl, double& xr, double& yb, double& yt) {
  cj = key / P.celly;
  ci = key - cj * P.celly;
  xl = (double)(cj - 1) * P.cs;
  xr = (double)(cj)*P.cs;
  yb = (double)(P.celly - 2 - ci) * P.cs;
  yt = (double)(P.celly - 1 - ci) * P.cs;
}

// quadrant of the particle in its cell (search.f90:328-388, porous.f90:133-178):
// returns 0 LL, 1 LR, 2 UR, 3 UL
__device__ __forceinline__ int quadrant(double x, double y, double xl, double xr, double yb,
                                        double yt) {
  const double xmid = xl + fabs((xl - xr) / 2);
  const double ymid = yb + fabs((yt - yb) / 2);
  if (x <= xmid && y <= ymid) return 0;
  if (x > xmid && y <= ymid) return 1;
  if (x > xmid && y > ymid) return 2;
  return 3;
}

// ---------------------------------------------------------------------
// deterministic block reduction: halving tree v[t] += v[t+s], s = 128..1.
// The first two levels go through LDS, the last six through wavefront
// shuffles inside wave 0; the result is identical to the LDS tree.
// ---------------------------------------------------------------------
__device__ __forceinline__ double block_tree(double val, double* lds) {
  const int t = threadIdx.x;
  lds[t] = val;
  __syncthreads();
  if (t < 128) lds[t] = lds[t] + lds[t + 128];
  __syncthreads();
  double v = 0.0;
  if (t < 64) {
    v = lds[t] + lds[t + 64];
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) v = v + __shfl_down(v, s, 64);
  }
  __syncthreads();
  return v;  // valid in thread 0
}

// block_tree of N values at once, each summed in block_tree's order (results
// in thread 0's v); lds holds N * RED_BLOCK doubles
template <int N>
__device__ __forceinline__ void block_tree_n(double (&v)[N], double* lds) {
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < N; k++) lds[k * RED_BLOCK + t] = v[k];
  __syncthreads();
  if (t < 128) {
#pragma unroll
    for (int k = 0; k < N; k++)
      lds[k * RED_BLOCK + t] = lds[k * RED_BLOCK + t] + lds[k * RED_BLOCK + t + 128];
  }
  __syncthreads();
  if (t < 64) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      double a = lds[k * RED_BLOCK + t] + lds[k * RED_BLOCK + t + 64];
#pragma unroll
      for (int s = 32; s >= 1; s >>= 1) a = a + __shfl_down(a, s, 64);
      v[k] = a;
    }
  }
  __syncthreads();
}

// maxima over a block of two non-negative double bit patterns (valid in
// thread 0); every thread of the block must call it
__device__ __forceinline__ void block_max2_u64(unsigned long long& a, unsigned long long& b) {
  __shared__ unsigned long long wa[RED_BLOCK / 64], wb[RED_BLOCK / 64];
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) {
    const unsigned long long oa = __shfl_xor(a, s, 64), ob = __shfl_xor(b, s, 64);
    a = (oa > a) ? oa : a;
    b = (ob > b) ? ob : b;
  }
  if ((threadIdx.x & 63) == 0) {
    wa[threadIdx.x >> 6] = a;
    wb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      a = (wa[w] > a) ? wa[w] : a;
      b = (wb[w] > b) ? wb[w] : b;
    }
  }
  __syncthreads();
}

// sum over a block (valid in thread 0); every thread of the block must call it
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v) {
  __shared__ unsigned long long ws[RED_BLOCK / 64];
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) v += ws[w];
  __syncthreads();
  return v;
}

// Final reduction of nb block partials in the fixed order: lane t adds
// partials t, t+256, ... sequentially, then the halving tree.  p1 (uniform
// over the block, or null) is a second partial array summed the same way;
// each lane keeps eight loads of each array in flight.
__device__ __forceinline__ void final_tree2(const double* __restrict__ p0,
                                            const double* __restrict__ p1, int nb, double* lds,
                                            double& s0, double& s1) {
  double a0 = 0.0, a1 = 0.0;
  for (int k0 = threadIdx.x; k0 < nb; k0 += 8 * RED_BLOCK) {
    double v0[8], v1[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int k = k0 + u * RED_BLOCK;
      v0[u] = (k < nb) ? p0[k] : 0.0;
      v1[u] = (p1 && k < nb) ? p1[k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (k0 + u * RED_BLOCK < nb) {
        a0 = a0 + v0[u];
        a1 = a1 + v1[u];
      }
  }
  s0 = block_tree(a0, lds);
  s1 = p1 ? block_tree(a1, lds) : 0.0;
}

// final_tree2 over N partial arrays p + k * stride (results in thread 0's s);
// lds holds N * RED_BLOCK doubles
template <int N>
__device__ __forceinline__ void final_tree_n(const double* __restrict__ p, int stride, int nb,
                                             double* lds, double (&s)[N]) {
#pragma unroll
  for (int k = 0; k < N; k++) s[k] = 0.0;
  for (int k0 = threadIdx.x; k0 < nb; k0 += 8 * RED_BLOCK) {
    double v[N][8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int j = k0 + u * RED_BLOCK;
#pragma unroll
      for (int k = 0; k < N; k++) v[k][u] = (j < nb) ? p[k * stride + j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (k0 + u * RED_BLOCK < nb) {
#pragma unroll
        for (int k = 0; k < N; k++) s[k] = s[k] + v[k][u];
      }
  }
  block_tree_n<N>(s, lds);
}
