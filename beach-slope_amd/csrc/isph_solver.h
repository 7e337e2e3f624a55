// isph_solver.h -- PPE solve kernels of the FP64 path: BiCGStab in the update
// order of the reference's bicgstab (solver.f90:182-213) with y = A x in the
// row order of spmv2 (solver.f90:84-127).
//
// On a breakdown below the iteration cap the host restarts from the true
// residual (k_bicg_init, FIN_RESTART), at most max_outer times.
//
// Each BiCGStab iteration is four row kernels, three of which write per-block
// partial dot products (block_tree over 256 rows, in row-block order whatever
// workgroup computes the block); a one-block finalize kernel then sums the
// partials in the fixed order (final_tree2) and forms alpha, omega or beta on
// the device.  The kernel boundary makes the
// partials visible, so no device-scope fence is needed inside the row
// kernels.  The host only polls the `done` flag every few iterations.
#pragma once
#include "isph_device.h"

// y_i = sum_k A_ik x_col(k), sequential in row order (spmv2)
__device__ __forceinline__ double row_spmv(const Dev& D, const double* __restrict__ val, int i,
                                           const double* __restrict__ x) {
  return ell_row_sum<ELL_U, double>(D.acol, val, D.cap, i, D.alen[i],
                                    [&](int j) { return x[j]; });
}

// ---- FP64 path ----
// Row kernels run on the XCD-contiguous row blocks of row_block(); block vb
// writes partial vb (and nblk + vb for a second dot product).

// r = b - A x ; rh = r ; p = r ; partial (r.r)
__global__ void __launch_bounds__(256) k_bicg_init(Dev D) {
  __shared__ double lds[RED_BLOCK];
  const int vb = row_block();
  if (vb >= D.nblk) return;
  const int i = vb * RED_BLOCK + threadIdx.x;
  const int n = D.sc->n;
  double prod = 0.0;
  if (i < n) {
    const double ax = row_spmv(D, D.aval, i, D.sol);
    const double r = D.b[i] - ax;
    D.r[i] = r;
    D.rh[i] = r;
    D.pp[i] = r;
    prod = r * r;
  }
  const double s = block_tree(prod, lds);
  if (threadIdx.x == 0) D.part[vb] = s;
}

// t1 = A p ; alpha = rho / (t1 . rh)
__global__ void __launch_bounds__(256) k_bicg_1(Dev D) {
  __shared__ double lds[RED_BLOCK];
  if (D.sc->done) return;
  const int vb = row_block();
  if (vb >= D.nblk) return;
  const int i = vb * RED_BLOCK + threadIdx.x;
  const int n = D.sc->n;
  double prod = 0.0;
  if (i < n) {
    const double t = row_spmv(D, D.aval, i, D.pp);
    D.t1[i] = t;
    prod = t * D.rh[i];
  }
  const double s = block_tree(prod, lds);
  if (threadIdx.x == 0) D.part[vb] = s;
}

// s = r - alpha t1 ; t2 = A s (s at the neighbours formed on the fly) ;
// omega = (t2 . s) / (t2 . t2)
__global__ void __launch_bounds__(256) k_bicg_2(Dev D) {
  __shared__ double lds[RED_BLOCK];
  if (D.sc->done) return;
  const int vb = row_block();
  if (vb >= D.nblk) return;
  const int i = vb * RED_BLOCK + threadIdx.x;
  const int n = D.sc->n;
  const double alpha = D.sc->alpha;
  double p1 = 0.0, p2 = 0.0;
  if (i < n) {
    const double* __restrict__ r = D.r;
    const double* __restrict__ t1 = D.t1;
    const double acc = ell_row_sum<ELL_U, double>(
        D.acol, D.aval, D.cap, i, D.alen[i], [&](int j) { return r[j] - alpha * t1[j]; });
    const double si = D.r[i] - alpha * D.t1[i];
    D.sv[i] = si;
    D.t2[i] = acc;
    p1 = acc * si;
    p2 = acc * acc;
  }
  const double s1 = block_tree(p1, lds);
  const double s2 = block_tree(p2, lds);
  if (threadIdx.x == 0) {
    D.part[vb] = s1;
    D.part[D.nblk + vb] = s2;
  }
}

// x = (x + alpha p) + omega s ; r = s - omega t2 ; rho' = r . rh ; rr = r . r
__global__ void __launch_bounds__(256) k_bicg_3(Dev D) {
  __shared__ double lds[RED_BLOCK];
  if (D.sc->done) return;
  const int vb = row_block();
  if (vb >= D.nblk) return;
  const int i = vb * RED_BLOCK + threadIdx.x;
  const int n = D.sc->n;
  const double alpha = D.sc->alpha, omega = D.sc->omega;
  double p1 = 0.0, p2 = 0.0;
  if (i < n) {
    const double si = D.sv[i];
    D.sol[i] = (D.sol[i] + alpha * D.pp[i]) + omega * si;
    const double r = si - omega * D.t2[i];
    D.r[i] = r;
    p1 = r * D.rh[i];
    p2 = r * r;
  }
  const double s1 = block_tree(p1, lds);
  const double s2 = block_tree(p2, lds);
  if (threadIdx.x == 0) {
    D.part[vb] = s1;
    D.part[D.nblk + vb] = s2;
  }
}

// p = r + beta (p - omega t1)
__global__ void k_bicg_4(Dev D) {
  if (D.sc->done) return;
  const int vb = row_block();
  if (vb >= D.nblk) return;
  const int i = vb * RED_BLOCK + threadIdx.x;
  if (i >= D.sc->n) return;
  const double beta = D.sc->beta, omega = D.sc->omega;
  D.pp[i] = D.r[i] + beta * (D.pp[i] - omega * D.t1[i]);
}

// generic dot (a . b) into *out through the fixed tree (used for ||b||^2)
__global__ void __launch_bounds__(256) k_dot(Dev D, const double* a, const double* b) {
  __shared__ double lds[RED_BLOCK];
  const int vb = row_block();
  if (vb >= D.nblk) return;
  const int i = vb * RED_BLOCK + threadIdx.x;
  const int n = D.sc->n;
  const double prod = (i < n) ? a[i] * b[i] : 0.0;
  const double s = block_tree(prod, lds);
  if (threadIdx.x == 0) D.part[vb] = s;
}

// pressure scatter (isph.f90:468-486): the solution vector is indexed by row
__global__ void k_copy_sol_to_p(Dev D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < D.sc->n) D.p[i] = D.sol[i];
}

// ---------------------------------------------------------------------
// one-block finalize kernels (256 threads) over the partials of the
// preceding row kernel
// ---------------------------------------------------------------------
enum {
  FIN_INIT = 0,     // rho = rr = sum ; convergence test
  FIN_ALPHA = 1,    // alpha = rho / sum
  FIN_OMEGA = 2,    // omega = sum0 / sum1
  FIN_BETA = 3,     // rho' = sum0, rr = sum1 ; iteration bookkeeping, beta
  FIN_DOT = 4,      // *out = sum
  FIN_RESTART = 5   // as FIN_INIT after a breakdown, keeping the iteration count
};

__global__ void __launch_bounds__(256) k_finalize(Dev D, int op, int nb, double* out) {
  __shared__ double lds[RED_BLOCK];
  DevScalars* S = D.sc;
  const bool skip = (op == FIN_ALPHA || op == FIN_OMEGA || op == FIN_BETA) && S->done;
  if (skip) return;
  double s0, s1;
  final_tree2(D.part, (op == FIN_OMEGA || op == FIN_BETA) ? D.part + nb : nullptr, nb, lds, s0,
              s1);
  if (threadIdx.x != 0) return;
  const DevParams& P = D.P;
  switch (op) {
    case FIN_INIT:
    case FIN_RESTART:
      S->rho = s0;
      S->rr = s0;
      if (op == FIN_INIT) S->iters = 0;
      S->brk = 0;
      S->conv = 0;
      S->done = 0;
      S->rel = sqrt(s0) / S->bnorm;
      if (sqrt(s0) <= P.tol * S->bnorm) {
        S->conv = 1;
        S->done = 1;
      }
      break;
    case FIN_ALPHA:
      S->alpha = S->rho / s0;
      break;
    case FIN_OMEGA:
      S->omega = s0 / s1;
      break;
    case FIN_BETA: {
      const double alpha = S->alpha, omega = S->omega, rhon = s0, rr = s1;
      S->iters = S->iters + 1;
      S->rr = rr;
      S->rel = sqrt(rr) / S->bnorm;
      if (sqrt(rr) <= P.tol * S->bnorm) {
        S->conv = 1;
        S->done = 1;
      } else if (!(isfinite(alpha) && isfinite(omega)) || omega == 0.0 || rhon == 0.0 ||
                 !isfinite(rhon)) {
        S->brk = 1;
        S->done = 1;
      } else {
        S->beta = (rhon / S->rho) * (alpha / omega);
        S->rho = rhon;
        if (S->iters >= P.max_iter) S->done = 1;
      }
      break;
    }
    case FIN_DOT:
      *out = s0;
      break;
  }
}
