// isph_wcsph.h -- the weakly compressible variant (wcsph.f90:13-154): Tait
// equation of state both ways, continuity with the delta-SPH diffusion term
// taken relative to the hydrostatic density, and the density update.  The
// reference module is dead and does not compile; these kernels follow its
// equations and expression order (see k_calmasscons for the one change).
// pow is the device library's: results agree with a host restatement to a
// few ulps, not bit for bit.
#pragma once
#include "isph_kernels.h"

// density of the hydrostatic pressure at height y (wcsph.f90:87-91)
__device__ __forceinline__ double d_rho_hyd(const DevParams& P, double y) {
  const double rho = P.rho;
  return rho * pow(((((P.swl - y) * rho) * fabs(P.g)) * P.gam) / (rho * (P.c0 * P.c0)) + 1.0,
                   1.0 / P.gam);
}

// presstoden, wcsph.f90:13-33 (every particle)
__global__ void k_presstoden(Dev D) {
  PIDX();
  if (i >= n) return;
  const DevParams& P = D.P;
  D.rho[i] = P.rho * pow(((D.p[i] * P.gam) / (P.rho * (P.c0 * P.c0))) + 1.0, 1.0 / P.gam);
}

// dentopress, wcsph.f90:35-63 (fluid; a negative pressure is clipped to 0
// and the density reset to rho)
__global__ void k_dentopress(Dev D) {
  PIDX();
  if (i >= n || D.tid[i] <= 2) return;
  const DevParams& P = D.P;
  double p = ((P.rho * (P.c0 * P.c0)) * (pow(D.rho[i] / P.rho, P.gam) - 1.0)) / P.gam;
  if (p < 0.0) {
    p = 0.0;
    D.rho[i] = P.rho;
  }
  D.p[i] = p;
}

// calmasscons, wcsph.f90:65-130: tden = rho + dt * sum over neighbours of
//   -m_j (u_j - u_i) . gradW~ (rho_i / rho_j)
//   + delt 2h c0 (rho_j - rho_i - (rhoH_j - rhoH_i)) (x_j - x_i) . gradW~ (m_j / rho_j) / (r^2 + lam)
// per component, gradW~ being the corrected gradient with respect to x_i.
// Two departures from the routine as written (it is never called there):
// its y-component sum shares the variable that holds the hydrostatic
// difference, so that sum is overwritten at every neighbour (the y-sum has
// its own accumulator here); and it takes (x_i - x_j), which with the
// gradient with respect to x_i makes the delta term anti-diffusive, where the
// delta-SPH scheme (Molteni & Colagrossi 2009) takes (x_j - x_i).
__global__ void k_calmasscons(Dev D) {
  PIDX();
  if (i >= n || D.tid[i] <= 2) return;
  const DevParams& P = D.P;
  const int cap = D.cap, cnt = D.ncnt[i];
  const double c1 = D.cf1[i], c2 = D.cf2[i], c3 = D.cf3[i], c4 = D.cf4[i];
  const double xn = D.xnb[i], yn = D.ynb[i];
  const double xi = D.x[i], yi = D.y[i], ui = D.u[i], vi = D.v[i], ri = D.rho[i];
  const double rhi = d_rho_hyd(P, yi);
  const double h = P.h, lam = P.lam, c0 = P.c0, delt = P.delt;
  double tx = 0.0, ty = 0.0;
  for (int k = 0; k < cnt; k++) {
    const int j = D.nbr[k * cap + i];
    const Klt z = d_klt(xn - D.xnb[j], yn - D.ynb[j], P);
    const double gx = c1 * z.k1 + c2 * z.k2, gy = c3 * z.k1 + c4 * z.k2;
    const double hyd = d_rho_hyd(P, D.y[j]) - rhi;
    const double mj = D.m[j], rj = D.rho[j];
    const double vj = mj / rj;
    const double ratio = ri / rj;
    const double dif = (rj - ri) - hyd;
    const double den = z.r * z.r + lam;
    tx = (tx - ((mj * (D.u[j] - ui)) * gx) * ratio) +
         ((((((delt * 2) * h) * c0) * dif) * (D.x[j] - xi)) * gx) * vj / den;
    ty = (ty - ((mj * (D.v[j] - vi)) * gy) * ratio) +
         ((((((delt * 2) * h) * c0) * dif) * (D.y[j] - yi)) * gy) * vj / den;
  }
  D.tden[i] = ri + (tx + ty) * D.sc->dt;
}

// updateden, wcsph.f90:132-152
__global__ void k_updateden(Dev D) {
  PIDX();
  if (i >= n || D.tid[i] <= 2) return;
  D.rho[i] = D.tden[i];
}
