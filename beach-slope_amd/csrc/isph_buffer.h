// isph_buffer.h -- inflow/outflow buffer zones: remove_buffer1/2 then
// insert_buffer1/2 (internal.f90:447-586), as one counting-sort pass.
//
// In each entry cell of a zone the particles that stay keep their order, the
// buffer particles past the zone's removal coordinate are dropped, and the
// zone's template particles of that cell are appended in template order;
// other cells keep their lists.  So the new list of a cell is (kept in order,
// then inserted in template order): one thread per cell ranks its kept
// particles, a scan gives the new cell starts, and two scatters write the
// kept and the inserted particles into the ping-pong state.
#pragma once
#include "isph_kernels.h"

struct BufZone {
  int zone;                        // 1: drop buffer particles with x >= rv; 2: with x <= rv
  int nt;                          // template particles
  const double *tx, *ty, *tu, *tv, *tp;
  const int32_t* tcell;            // template particle's cell
  const int32_t* trank;            // its rank among the template particles of that cell
  const int32_t* tcnt;             // per cell: template particles (0 outside the entry cells)
  const uint8_t* entry;            // per cell: 1 for an entry cell of the zone
  double mass, rho, vol0, c, rv;   // values an inserted particle takes
};

// one thread per cell: rank of each kept particle within its cell (-1 when
// dropped); new count (kept + inserted) in bincnt[k], kept count in
// bincnt[ncell + 1 + k]
__global__ void k_buf_rank(Dev D, BufZone Z, int ncell) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncell) return;
  const int i0 = D.cstart[k], i1 = D.cstart[k + 1];
  const bool ent = Z.entry[k] != 0;
  int kept = 0;
  for (int i = i0; i < i1; i++) {
    bool drop = false;
    if (ent && (D.flags[i] & F_BUFFER)) drop = (Z.zone == 1) ? (D.x[i] >= Z.rv) : (D.x[i] <= Z.rv);
    D.prank[i] = drop ? -1 : kept;
    if (!drop) kept++;
  }
  D.bincnt[k] = kept + Z.tcnt[k];
  D.bincnt[ncell + 1 + k] = kept;
}

// kept particles to their new slots; porosity moves with the particle (into
// the solver vector r, free between solves)
__global__ void k_buf_scatter_kept(Dev D) {
  PIDX();
  if (i >= n) return;
  const int rk = D.prank[i];
  if (rk < 0) return;
  const int d = D.binstart[D.cell[i]] + rk;
  D.x2[d] = D.x[i]; D.y2[d] = D.y[i]; D.u2[d] = D.u[i]; D.v2[d] = D.v[i];
  D.us2[d] = D.us[i]; D.vs2[d] = D.vs[i]; D.xs2[d] = D.xs[i]; D.ys2[d] = D.ys[i];
  D.p2[d] = D.p[i]; D.m2[d] = D.m[i]; D.rho2[d] = D.rho[i]; D.rho02[d] = D.rho0[i];
  D.vol02[d] = D.vol0[i]; D.c2[d] = D.c[i];
  D.tid2[d] = D.tid[i]; D.flags2[d] = D.flags[i]; D.vic2[d] = D.vic[i];
  D.cell2[d] = D.cell[i];
  D.r[d] = D.por[i];
}

// template particles after the kept ones of their cell (insert_buffer1/2):
// position, velocity and pressure from the template; tid 3, buffer set, the
// zone's mass, density (also the reference density), volume, concentration;
// intermediate values zero; porosity 1
__global__ void k_buf_scatter_new(Dev D, BufZone Z, int ncell) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Z.nt) return;
  const int k = Z.tcell[t];
  const int d = D.binstart[k] + D.bincnt[ncell + 1 + k] + Z.trank[t];
  D.x2[d] = Z.tx[t]; D.y2[d] = Z.ty[t]; D.u2[d] = Z.tu[t]; D.v2[d] = Z.tv[t];
  D.us2[d] = 0.0; D.vs2[d] = 0.0; D.xs2[d] = 0.0; D.ys2[d] = 0.0;
  D.p2[d] = Z.tp[t]; D.m2[d] = Z.mass; D.rho2[d] = Z.rho; D.rho02[d] = Z.rho;
  D.vol02[d] = Z.vol0; D.c2[d] = Z.c;
  D.tid2[d] = TID_FLUID; D.flags2[d] = F_BUFFER; D.vic2[d] = 0;
  D.cell2[d] = k;
  D.r[d] = 1.0;
}

// new cell starts and count; porosity back from r; solution rows beyond the
// old count start from 0
__global__ void k_buf_finish(Dev D, int ncell, int nt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int n_old = D.sc->n, n_new = D.binstart[ncell];
  if (c <= ncell) D.cstart[c] = D.binstart[c];
  if (c < n_new) D.por[c] = D.r[c];
  if (c < nt && n_old + c < n_new) D.sol[n_old + c] = 0.0;
}

__global__ void k_buf_set_n(Dev D, int ncell) {
  if (blockIdx.x == 0 && threadIdx.x == 0) D.sc->n = D.binstart[ncell];
}
