/*
 * isph_oracle.c -- CPU restatement of the reference's per-step ISPH pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it,
 * and only to check (or time, as a baseline) the HIP path.  The product path
 * never calls it.
 *
 * It restates the reference Fortran (Sinchanrc/Beach-Slope) expression by
 * expression, evaluating every expression left to right as written in the
 * Fortran source, with real4 literals rounded to float as Fortran does
 * (kernel.f90:61-62, functions.f90:324, integrator.f90:498/581/624,
 * scalar.f90:246/322).  Build with -O2 -ffp-contract=off (no FMA, no
 * reassociation).  Where the reference is data-structure specific the oracle
 * follows it literally:
 *   - rebinning walks each cell's slots and builds the reference's nine
 *     transfer lists, then concatenates walls, stayers and incomers in the
 *     order of integrator.f90:324-386;
 *   - the neighbour search walks the quadrant 2x2 block cell by cell
 *     (search.f90:310-399) and STORES r and grad W per neighbour
 *     (search.f90:61-74), as the reference's verlet list does;
 *   - the PPE rows are filled as padded rows exactly as isph.f90:123-456
 *     (val(1)=1 and col=row initialisation, ins_pt/dia_pt, Jacobi) and then
 *     written in canonical order (columns ascending, the diagonal before the
 *     row's untouched slot-1 value, explicit zeros dropped).
 * Deliberate, documented deviations: BiCGStab stops on the relative residual
 * ||r||/||b|| (the reference's unused bicgstab stops on an update test,
 * solver.f90:233-242, and its live solver is MKL FGMRES at 1e-3,
 * solver.f90:306-350); dot products use a fixed blocked tree (see dot_tree);
 * a zero shift stores 0 instead of the reference's 0/0 (part_shift.f90:89-90).
 *
 * Derived per-particle properties (the reference's per-slot `pplist`,
 * domain.f90:36) are not carried through a rebin: the oracle poisons them with
 * NaN after every rebin, so a read of a stale property before it is
 * recomputed shows up as NaN in every parity comparison.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/isph_abi.h"

/* Planted mistakes for the pin tests (tests/test_oracle_pins.py): a build
 * with -DOR_MUTANT=k changes the one expression marked MUT(k) below, so a pin
 * can show that it fails on a plausible transcription error.  The default
 * build (OR_MUTANT = 0) has none of them. */
#ifndef OR_MUTANT
#define OR_MUTANT 0
#endif
#define MUT(k) (OR_MUTANT == (k))

#define KCAP ISPH_NBR_CAP
#define RCAP (ISPH_NBR_CAP + 2)
#define RED_BLOCK 256
#define OR_PI 3.14159265358979323846 /* M_PI of <math.h> */

/* an inflow/outflow buffer zone (internal.f90:447-586): the template
 * particles buffer1/buffer2, their cells (the entry cells), the values an
 * inserted particle takes, and the insertion clock of main.f90:217-235 */
typedef struct ozone {
  int64_t nt;
  double *tx, *ty, *tu, *tv, *tp;
  int32_t *tcell;
  uint8_t *entry;               /* per cell: an entry cell of this zone */
  double mass, rho, vol0, c, rv, interval, time_shift;
  int32_t iter;                 /* next insertion is at iter * interval */
  int set;
} ozone;

typedef struct ostate {
  isph_params P;
  ozone zone[2];
  /* WCSPH (wcsph.f90): sound speed co, exponent t_gam, delta-SPH coefficient
   * delt, still-water level of the hydrostatic density; in_step switches
   * the fused step to the weakly compressible form */
  double c0, gam, delt, swl;
  int wcsph;
  /* piston wavemaker: stroke, period, ramp time, start time; the paddle's
   * rest positions live in xs */
  double wv_stroke, wv_period, wv_ramp, wv_t0;
  int wave;
  double *tden;
  int64_t n, cap, ns, scap;
  int32_t ncell;
  /* particle fields in cell-sorted order (the reference's plist) */
  double *x, *y, *u, *v, *us, *vs, *xs, *ys, *p, *m, *rho, *rho0, *vol0, *c;
  int32_t *tid, *cell;
  uint8_t *flags;
  /* derived properties (the reference's pplist) */
  double *por, *resist, *gradvx, *coff, *nut, *tke, *varts, *cdiff, *vel;
  /* verlet lists: count, neighbour index, stored r and grad W (klt) */
  int32_t *ncnt, *nbr;
  double *nr, *klt1, *klt2;
  int32_t *cstart;
  /* solids */
  double *sx, *sy, *sm, *srho;
  int32_t *scell, *sstart;
  /* PPE: canonical rows, rhs, persistent solution (fsol by row) */
  int32_t *alen, *acol;
  double *aval, *b, *sol;
  int assembled;
  /* scalars */
  double t, dt, dtsol, umax, numax;
  /* per cell: max |vel| and max |nu_t| over the cell's slots at the last
   * comp_vel that found the cell non-empty (dpcell%maxvel / maxeddy,
   * integrator.f90:831-841); timestep reduces them over every cell */
  double *cmaxvel, *cmaxeddy;
  int32_t solsteps;
  isph_solve_stats last;
  /* scratch */
  double *w1, *w2, *w3, *w4, *w5, *w6, *w7, *w8;
  int32_t *perm;
  uint8_t *cls;
} ostate;

/* ------------------------------------------------------------------ */
/* kernel.f90 / functions.f90 restatement                               */
/* ------------------------------------------------------------------ */

/* 28.0*3.14 and 3.0*3.14 are real4 products in kernel.f90:61-62. */
static double k1c(double h) { return 10.0 / ((double)(28.0f * 3.14f) * (h * h)); }
static double k2c(double h) { return 1.0 / ((double)(3.0f * 3.14f) * (h * h)); }

/* Wab, kernel.f90:34-66 (hWab, kernel.f90:164-184, is identical). */
static double o_wab(double r, double h) {
  double t1, t2;
  if (r < h) {
    t1 = (4.0 - ((6.0 * (r * r)) / (h * h))) + ((3.0 * ((r * r) * r)) / ((h * h) * h));
    double q = r / h;
    t2 = (((q * q) * q) - ((6.0 * r) / h)) + 6.0;
  } else {
    double a = 2.0 - r / h;
    t1 = (a * a) * a;
    t2 = (a * a) * a;
  }
  return ((0.1 * t1) * k1c(h)) + (((1.0 - 0.1) * t2) * k2c(h));
}

/* Wo, kernel.f90:68-92. */
static double o_wo(double h) {
  return ((0.1 * 4.0) * k1c(h)) + (((1.0 - 0.1) * 6.0) * k2c(h));
}

/* Wabx(p1=j, p2=i, r, h), kernel.f90:94-127; d_ij = coordinate of p2 minus
 * coordinate of p1.  Waby (129-162), hWabx/hWaby (204-248) are identical. */
static double o_wabd(double d_ij, double r, double h) {
  double t1, t2;
  if (r < h) {
    t1 = (((9.0 * r) / ((h * h) * h)) - (12.0 / (h * h))) * d_ij;
    t2 = (((3.0 * r) / ((h * h) * h)) - (6.0 / (r * h))) * d_ij;
  } else {
    double a = 2.0 - r / h;
    t1 = ((3.0 / (r * h)) * (a * a)) * (-d_ij);
    t2 = ((3.0 / (r * h)) * (a * a)) * (-d_ij);
  }
  return ((0.1 * t1) * k1c(h)) + (((1.0 - 0.1) * t2) * k2c(h));
}

/* dist, functions.f90:28-42 */
static double o_dist(double x1, double y1, double x2, double y2) {
  double dx = x1 - x2, dy = y1 - y2;
  return sqrt((dx * dx) + (dy * dy));
}

/* invertmat2D, functions.f90:62-96 (1e-5 is a real4 literal). */
static void o_inv2(double *a) {
  double det = a[0] * a[3] - a[1] * a[2];
  if ((fabs(det) >= (double)1e-5f) && !isnan(det) && isfinite(det)) {
    double r0 = a[3] / det, r1 = (-a[1]) / det, r2 = (-a[2]) / det, r3 = a[0] / det;
    a[0] = r0; a[1] = r1; a[2] = r2; a[3] = r3;
  } else {
    a[0] = 1.0; a[1] = 0.0; a[2] = 0.0; a[3] = 1.0;
  }
}

/* vis, porous.f90:12-30 and integrator.f90:12-30 */
static double o_vis(double oden) {
  double sp = oden - 1000.0;
  return ((1.002 + (0.001652 * sp)) + (0.0000083 * (sp * sp))) * 0.001;
}

/* Fortran MIN of two reals as used here (NaN in the first argument kept). */
static double o_min(double a, double b) { return (b < a) ? b : a; }

/* ------------------------------------------------------------------ */
/* cell geometry, setup.f90:114-127                                     */
/* ------------------------------------------------------------------ */
static double cs_of(const ostate *s) { return 4.0 * s->P.h; }
static double xleft(const ostate *s, int cj) { return (double)(cj - 1) * cs_of(s); }
static double xright(const ostate *s, int cj) { return (double)(cj)*cs_of(s); }
static double ybot(const ostate *s, int ci) { return (double)(s->P.celly - 2 - ci) * cs_of(s); }
static double ytop(const ostate *s, int ci) { return (double)(s->P.celly - 1 - ci) * cs_of(s); }
static int key_of(const ostate *s, int ci, int cj) { return cj * s->P.celly + ci; }
static int interior(const ostate *s, int ci, int cj) {
  return ci >= 1 && ci <= s->P.celly - 2 && cj >= 1 && cj <= s->P.cellx - 2;
}

/* ------------------------------------------------------------------ */
/* allocation                                                           */
/* ------------------------------------------------------------------ */
#define ALLOC(p, cnt) (p) = calloc((size_t)(cnt) > 0 ? (size_t)(cnt) : 1, sizeof(*(p)))

ostate *or_create(const isph_params *P, int64_t cap, int64_t scap) {
  ostate *s = calloc(1, sizeof(ostate));
  s->P = *P;
  s->cap = cap;
  s->scap = scap;
  s->ncell = P->cellx * P->celly;
  double **dp[] = {&s->x, &s->y, &s->u, &s->v, &s->us, &s->vs, &s->xs, &s->ys, &s->p, &s->m,
                   &s->rho, &s->rho0, &s->vol0, &s->c, &s->por, &s->resist, &s->gradvx,
                   &s->nut, &s->tke, &s->varts, &s->cdiff, &s->vel, &s->b, &s->sol,
                   &s->w1, &s->w2, &s->w3, &s->w4, &s->w5, &s->w6, &s->w7, &s->w8,
                   &s->tden};
  for (size_t k = 0; k < sizeof(dp) / sizeof(dp[0]); k++) ALLOC(*dp[k], cap);
  ALLOC(s->coff, 4 * cap);
  ALLOC(s->tid, cap);
  ALLOC(s->cell, cap);
  ALLOC(s->flags, cap);
  ALLOC(s->perm, cap);
  ALLOC(s->cls, cap);
  ALLOC(s->ncnt, cap);
  ALLOC(s->nbr, cap * KCAP);
  ALLOC(s->nr, cap * KCAP);
  ALLOC(s->klt1, cap * KCAP);
  ALLOC(s->klt2, cap * KCAP);
  ALLOC(s->alen, cap);
  ALLOC(s->acol, cap * RCAP);
  ALLOC(s->aval, cap * RCAP);
  ALLOC(s->cstart, s->ncell + 1);
  ALLOC(s->sx, scap);
  ALLOC(s->sy, scap);
  ALLOC(s->sm, scap);
  ALLOC(s->srho, scap);
  ALLOC(s->scell, scap);
  ALLOC(s->sstart, s->ncell + 1);
  ALLOC(s->cmaxvel, s->ncell);
  ALLOC(s->cmaxeddy, s->ncell);
  s->dt = P->dt;
  s->solsteps = 1;
  s->dtsol = P->dt;
  return s;
}

void or_destroy(ostate *s) {
  if (!s) return;
  void *ptrs[] = {s->x, s->y, s->u, s->v, s->us, s->vs, s->xs, s->ys, s->p, s->m, s->rho,
                  s->rho0, s->vol0, s->c, s->por, s->resist, s->gradvx, s->nut, s->tke,
                  s->varts, s->cdiff, s->vel, s->b, s->sol, s->w1, s->w2, s->w3, s->w4,
                  s->w5, s->w6, s->w7, s->w8, s->coff, s->tid, s->cell, s->flags, s->perm, s->cls, s->ncnt,
                  s->nbr, s->nr, s->klt1, s->klt2, s->alen, s->acol, s->aval,
                  s->cstart, s->sx, s->sy, s->sm, s->srho, s->scell, s->sstart, s->tden,
                  s->cmaxvel, s->cmaxeddy};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
  for (int z = 0; z < 2; z++) {
    ozone *Z = &s->zone[z];
    free(Z->tx); free(Z->ty); free(Z->tu); free(Z->tv); free(Z->tp); free(Z->tcell);
    free(Z->entry);
  }
  free(s);
}

int64_t or_n(const ostate *s) { return s->n; }
int64_t or_ns(const ostate *s) { return s->ns; }
void or_set_params(ostate *s, const isph_params *P) { s->P = *P; }

/* containing cell with half-open bounds (internal.f90:80-83); -1 outside */
static int locate(const ostate *s, double x, double y, int *ci_out, int *cj_out) {
  double cs = cs_of(s);
  if (!(x == x) || !(y == y)) return -1;
  int cj = (int)floor(x / cs) + 1;
  int ci = s->P.celly - 2 - (int)floor(y / cs);
  for (int it = 0; it < 3; it++) {
    if (cj < 0 || cj >= s->P.cellx) break;
    if (x < xleft(s, cj)) cj--;
    else if (x >= xright(s, cj)) cj++;
  }
  for (int it = 0; it < 3; it++) {
    if (ci < 0 || ci >= s->P.celly) break;
    if (y < ybot(s, ci)) ci++;
    else if (y >= ytop(s, ci)) ci--;
  }
  if (cj < 0 || cj >= s->P.cellx || ci < 0 || ci >= s->P.celly) return -1;
  if (!(x >= xleft(s, cj) && x < xright(s, cj) && y >= ybot(s, ci) && y < ytop(s, ci)))
    return -1;
  if (!interior(s, ci, cj)) return -1;
  *ci_out = ci;
  *cj_out = cj;
  return key_of(s, ci, cj);
}

static void poison_props(ostate *s) {
  const double nan = NAN;
  for (int64_t i = 0; i < s->n; i++) {
    int wall = s->tid[i] != ISPH_TID_FLUID;
    s->por[i] = wall ? 1.0 : nan;
    s->resist[i] = wall ? 0.0 : nan;
    s->gradvx[i] = nan;
    for (int k = 0; k < 4; k++) s->coff[4 * i + k] = nan;
    s->nut[i] = wall ? 0.0 : nan;
    s->tke[i] = wall ? 0.0 : nan;
    s->varts[i] = nan;
    s->cdiff[i] = nan;
    s->vel[i] = wall ? 0.0 : nan;
    s->ncnt[i] = -1;
  }
}

static void rebuild_starts(ostate *s) {
  for (int k = 0; k <= s->ncell; k++) s->cstart[k] = 0;
  for (int64_t i = 0; i < s->n; i++) s->cstart[s->cell[i] + 1]++;
  for (int k = 0; k < s->ncell; k++) s->cstart[k + 1] += s->cstart[k];
}

/* Initial binning: walls first then fluid in each cell, input order kept
 * (walls are generated into the cells first, boundary.f90:159-185, then the
 * fluid lattice is appended, internal.f90:80-91). */
int or_upload(ostate *s, int64_t n, const double *x, const double *y, const double *u,
              const double *v, const double *p, const double *m, const double *rho,
              const double *rho0, const double *vol0, const double *c, const int32_t *tid,
              const uint8_t *flags) {
  if (n > s->cap || n < 0) return ISPH_ECAP;
  int32_t *key = malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  int32_t *cnt = calloc((size_t)s->ncell * 2 + 1, sizeof(int32_t));
  for (int64_t i = 0; i < n; i++) {
    int ci, cj;
    int k = locate(s, x[i], y[i], &ci, &cj);
    if (k < 0 || (tid[i] == ISPH_TID_FLUID && (flags[i] & ISPH_F_PADDLE))) {
      free(key);
      free(cnt);
      return ISPH_EARG;
    }
    key[i] = 2 * k + (tid[i] == ISPH_TID_FLUID ? 1 : 0);
    cnt[key[i] + 1]++;
  }
  for (int k = 0; k < 2 * s->ncell; k++) cnt[k + 1] += cnt[k];
  for (int64_t i = 0; i < n; i++) {
    int64_t d = cnt[key[i]]++;
    s->x[d] = x[i]; s->y[d] = y[i]; s->u[d] = u[i]; s->v[d] = v[i]; s->p[d] = p[i];
    s->m[d] = m[i]; s->rho[d] = rho[i]; s->rho0[d] = rho0[i]; s->vol0[d] = vol0[i];
    s->c[d] = c[i]; s->tid[d] = tid[i]; s->flags[d] = flags[i];
    s->cell[d] = key[i] / 2;
    s->us[d] = 0.0; s->vs[d] = 0.0; s->xs[d] = 0.0; s->ys[d] = 0.0;
  }
  s->n = n;
  free(key);
  free(cnt);
  rebuild_starts(s);
  poison_props(s);
  /* fsol starts from the initial pressures (memory.f90:113-144) */
  for (int64_t i = 0; i < n; i++) s->sol[i] = s->p[i];
  s->assembled = 0;
  return 0;
}

int or_upload_solids(ostate *s, int64_t ns, const double *x, const double *y,
                     const double *m, const double *rho) {
  if (ns > s->scap || ns < 0) return ISPH_ECAP;
  int32_t *key = malloc(sizeof(int32_t) * (ns > 0 ? ns : 1));
  for (int k = 0; k <= s->ncell; k++) s->sstart[k] = 0;
  for (int64_t i = 0; i < ns; i++) {
    int ci, cj;
    int k = locate(s, x[i], y[i], &ci, &cj);
    if (k < 0) {
      free(key);
      return ISPH_EARG;
    }
    key[i] = k;
    s->sstart[k + 1]++;
  }
  for (int k = 0; k < s->ncell; k++) s->sstart[k + 1] += s->sstart[k];
  int32_t *fill = malloc(sizeof(int32_t) * (s->ncell + 1));
  memcpy(fill, s->sstart, sizeof(int32_t) * (s->ncell + 1));
  for (int64_t i = 0; i < ns; i++) {
    int64_t d = fill[key[i]]++;
    s->sx[d] = x[i]; s->sy[d] = y[i]; s->sm[d] = m[i]; s->srho[d] = rho[i];
    s->scell[d] = key[i];
  }
  s->ns = ns;
  free(key);
  free(fill);
  return 0;
}

/* ------------------------------------------------------------------ */
/* projection, isph.f90:17-47                                           */
/* ------------------------------------------------------------------ */
void or_projection(ostate *s) {
  const double dt = s->dt;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] == ISPH_TID_FLUID) {
      s->xs[i] = s->x[i];
      s->x[i] = MUT(14) ? s->x[i] + ((dt * s->u[i]) * s->por[i])
                        : s->x[i] + ((dt * s->u[i]) / s->por[i]);
      s->ys[i] = s->y[i];
      s->y[i] = MUT(14) ? s->y[i] + ((dt * s->v[i]) * s->por[i])
                        : s->y[i] + ((dt * s->v[i]) / s->por[i]);
    }
  }
}

/* ------------------------------------------------------------------ */
/* cellshift / cellshiftalt, integrator.f90:33-392                     */
/* ------------------------------------------------------------------ */
enum { C_WALL = 0, C_STAY, C_TN, C_TS, C_TW, C_TE, C_TSE, C_TNE, C_TSW, C_TNW, C_DROP = 15 };

static void gather_particles(ostate *s, const int32_t *src, int64_t nnew, const int32_t *newcell) {
  double **fd[] = {&s->x, &s->y, &s->u, &s->v, &s->us, &s->vs, &s->xs, &s->ys,
                   &s->p, &s->m, &s->rho, &s->rho0, &s->vol0, &s->c};
  for (size_t f = 0; f < sizeof(fd) / sizeof(fd[0]); f++) {
    double *a = *fd[f];
    for (int64_t d = 0; d < nnew; d++) s->w1[d] = a[src[d]];
    memcpy(a, s->w1, sizeof(double) * (size_t)nnew);
  }
  int32_t *ti = (int32_t *)s->w2;
  for (int64_t d = 0; d < nnew; d++) ti[d] = s->tid[src[d]];
  memcpy(s->tid, ti, sizeof(int32_t) * (size_t)nnew);
  uint8_t *fl = (uint8_t *)s->w2;
  for (int64_t d = 0; d < nnew; d++) fl[d] = s->flags[src[d]];
  memcpy(s->flags, fl, (size_t)nnew);
  memcpy(s->cell, newcell, sizeof(int32_t) * (size_t)nnew);
  s->n = nnew;
}

static void or_rebin(ostate *s, int alt) {
  const int celly = s->P.celly, cellx = s->P.cellx;
  /* classification pass: one transfer class per fluid particle */
  for (int64_t i = 0; i < s->n; i++) {
    /* walls keep their slots; wavemaker paddle walls move as fluid does */
    if (s->tid[i] != ISPH_TID_FLUID && !(s->flags[i] & ISPH_F_PADDLE)) {
      s->cls[i] = C_WALL;
      continue;
    }
    int key = s->cell[i], cj = key / celly, ci = key % celly;
    double x1 = s->x[i], y1 = s->y[i];
    double xl = xleft(s, cj), xr = xright(s, cj), yb = ybot(s, ci), yt = ytop(s, ci);
    int c = C_DROP;
    if ((x1 >= xl) && (x1 < xr)) {
      if ((y1 >= yb) && (y1 < yt)) c = C_STAY;
      else if (y1 < yb) c = C_TS;
      else if (y1 >= yt) c = C_TN;
    } else if (x1 >= xr) {
      if (y1 >= yt) c = C_TNE;
      else if ((y1 >= yb) && (y1 < yt)) c = C_TE;
      else if (y1 < yb) c = C_TSE;
    } else if (x1 < xl) {
      if (y1 >= yt) c = C_TNW;
      else if ((y1 >= yb) && (y1 < yt)) c = C_TW;
      else if (y1 < yb) c = C_TSW;
    }
    if (alt && c != C_DROP && s->tid[i] == ISPH_TID_FLUID) {
      int buf = (s->flags[i] & ISPH_F_BUFFER) != 0;
      if (!((x1 > s->P.xlcutoff) || buf || ((x1 <= s->P.xlcutoff) && (y1 <= s->P.ytcutoff))))
        c = C_DROP;
    }
    s->cls[i] = (uint8_t)c;
  }
  /* transfer pass: rebuild every interior cell in the reference's order */
  int64_t nnew = 0;
  int32_t *newcell = (int32_t *)s->w3;
  static const int order_cls[8] = {C_TN, C_TS, C_TW, C_TE, C_TSE, C_TNE, C_TSW, C_TNW};
  /* source cell offset (dci, dcj) of each incoming list */
  static const int src_di[8] = {+1, -1, 0, 0, -1, +1, -1, +1};
  static const int src_dj[8] = {0, 0, +1, -1, -1, -1, +1, +1};
  for (int cj = 0; cj < cellx; cj++) {
    for (int ci = 0; ci < celly; ci++) {
      if (!interior(s, ci, cj)) continue;
      int key = key_of(s, ci, cj);
      for (int32_t k = s->cstart[key]; k < s->cstart[key + 1]; k++)
        if (s->cls[k] == C_WALL) { s->perm[nnew] = k; newcell[nnew++] = key; }
      for (int32_t k = s->cstart[key]; k < s->cstart[key + 1]; k++)
        if (s->cls[k] == C_STAY) { s->perm[nnew] = k; newcell[nnew++] = key; }
      for (int l = 0; l < 8; l++) {
        int sci = ci + src_di[l], scj = cj + src_dj[l];
        if (!interior(s, sci, scj)) continue;
        int skey = key_of(s, sci, scj);
        for (int32_t k = s->cstart[skey]; k < s->cstart[skey + 1]; k++)
          if (s->cls[k] == order_cls[l]) { s->perm[nnew] = k; newcell[nnew++] = key; }
      }
    }
  }
  gather_particles(s, s->perm, nnew, newcell);
  rebuild_starts(s);
  poison_props(s);
  s->assembled = 0;
}

void or_cellshift(ostate *s) { or_rebin(s, 0); }
void or_cellshiftalt(ostate *s) { or_rebin(s, 1); }

/* ------------------------------------------------------------------ */
/* neighbour / nsearch2, search.f90:13-399                              */
/* ------------------------------------------------------------------ */
/* quadrant block of a particle in its cell: 4 cell keys in traversal order */
static void quad_block(const ostate *s, int64_t i, int blk[4]) {
  int key = s->cell[i], cj = key / s->P.celly, ci = key % s->P.celly;
  double xl = xleft(s, cj), xr = xright(s, cj), yb = ybot(s, ci), yt = ytop(s, ci);
  double xmid = xl + fabs((xl - xr) / 2), ymid = yb + fabs((yt - yb) / 2);
  double x = s->x[i], y = s->y[i];
  int di[4], dj[4];
  if (x <= xmid && y <= ymid) {        /* search.f90:335-338 */
    int a[4] = {0, 1, 0, 1}, b[4] = {-1, -1, 0, 0};
    memcpy(di, a, sizeof a); memcpy(dj, b, sizeof b);
  } else if (x > xmid && y <= ymid) {  /* search.f90:347-350 */
    int a[4] = {0, 1, 0, 1}, b[4] = {0, 0, 1, 1};
    memcpy(di, a, sizeof a); memcpy(dj, b, sizeof b);
  } else if (x > xmid && y > ymid) {   /* search.f90:359-362 */
    int a[4] = {-1, 0, -1, 0}, b[4] = {0, 0, 1, 1};
    memcpy(di, a, sizeof a); memcpy(dj, b, sizeof b);
  } else {                              /* search.f90:371-374 (and NaN) */
    int a[4] = {-1, 0, -1, 0}, b[4] = {-1, -1, 0, 0};
    memcpy(di, a, sizeof a); memcpy(dj, b, sizeof b);
  }
  for (int k = 0; k < 4; k++) blk[k] = key_of(s, ci + di[k], cj + dj[k]);
}

int or_neighbour(ostate *s) {
  const double h = s->P.h;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : err)
  for (int64_t i = 0; i < s->n; i++) {
    int blk[4];
    quad_block(s, i, blk);
    double res[4] = {0.0, 0.0, 0.0, 0.0};
    int cnt = 0;
    s->flags[i] &= (uint8_t)~ISPH_F_FREE;
    s->gradvx[i] = 0.0;
    for (int cb = 0; cb < 4; cb++) {
      for (int32_t j = s->cstart[blk[cb]]; j < s->cstart[blk[cb] + 1]; j++) {
        /* tid differ, or same tid and different matid: i.e. j is not i */
        if (!((s->tid[i] != s->tid[j]) || (j != i))) continue;
        double r = o_dist(s->x[i], s->y[i], s->x[j], s->y[j]);
        if (!(r < (2 * h))) continue;
        if (cnt >= KCAP) { err = 1; continue; }
        double k1 = o_wabd(s->x[i] - s->x[j], r, h);
        double k2 = o_wabd(s->y[i] - s->y[j], r, h);
        s->nbr[i * KCAP + cnt] = j;
        s->nr[i * KCAP + cnt] = r;
        s->klt1[i * KCAP + cnt] = k1;
        s->klt2[i * KCAP + cnt] = k2;
        cnt++;
        double r1 = (s->m[j] * (s->x[j] - s->x[i])) / s->rho[j];
        double r3 = (s->m[j] * (s->y[j] - s->y[i])) / s->rho[j];
        res[0] = res[0] + k1 * r1;
        res[1] = res[1] + k2 * r1;
        res[2] = res[2] + k1 * r3;
        res[3] = res[3] + k2 * r3;
      }
    }
    s->ncnt[i] = cnt;
    if (s->tid[i] == ISPH_TID_FLUID) {
      s->flags[i] &= (uint8_t)~(ISPH_F_FREE | ISPH_F_VICINITY);
      s->gradvx[i] = res[0] + res[3];
      o_inv2(res);
      for (int k = 0; k < 4; k++) s->coff[4 * i + k] = res[k];
      if ((s->gradvx[i] <= s->P.fs_thresh) && !(s->flags[i] & ISPH_F_BUFFER))
        s->flags[i] |= ISPH_F_FREE;
    }
  }
  return err ? ISPH_ECAP : 0;
}

/* ------------------------------------------------------------------ */
/* effpor / porsearch, porous.f90:32-224                                */
/* ------------------------------------------------------------------ */
void or_effpor(ostate *s) {
  const double h = s->P.h, por = s->P.por, lamk = s->P.lamk, d50 = s->P.d50;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    s->resist[i] = 0.0;
    if (s->tid[i] != ISPH_TID_FLUID) continue;
    s->por[i] = 1.0;
    int key = s->cell[i], cj = key / s->P.celly, ci = key % s->P.celly;
    double xl = xleft(s, cj), xr = xright(s, cj), yb = ybot(s, ci), yt = ytop(s, ci);
    double xmid = xl + fabs((xl - xr) / 2), ymid = yb + fabs((yt - yb) / 2);
    double x = s->x[i], y = s->y[i];
    int di[4] = {0, 0, 0, 0}, dj[4] = {0, 0, 0, 0};
    if (x <= xmid && y <= ymid) {         /* porous.f90:138-140 */
      di[1] = 0; dj[1] = MUT(5) ? 1 : -1; di[2] = 1; dj[2] = 0; di[3] = 1; dj[3] = MUT(5) ? 1 : -1;
    } else if (x > xmid && y <= ymid) {   /* porous.f90:147-149 */
      di[1] = 0; dj[1] = 1; di[2] = 1; dj[2] = 0; di[3] = 1; dj[3] = 1;
    } else if (x > xmid && y > ymid) {    /* porous.f90:156-158 */
      di[1] = -1; dj[1] = 1; di[2] = -1; dj[2] = 0; di[3] = 0; dj[3] = 1;
    } else {                               /* porous.f90:165-167 */
      di[1] = -1; dj[1] = 0; di[2] = -1; dj[2] = -1; di[3] = 0; dj[3] = -1;
    }
    double sf = 0.0;
    for (int cb = 0; cb < 4; cb++) {
      int k = key_of(s, ci + di[cb], cj + dj[cb]);
      for (int32_t q = s->sstart[k]; q < s->sstart[k + 1]; q++) {
        double r = o_dist(x, y, s->sx[q], s->sy[q]);
        if (r < (2 * h)) sf = sf + ((((1.0 - por * 1.0) * o_wab(r, h)) * s->sm[q]) / s->srho[q]);
      }
    }
    if (sf > (1.0 - por)) sf = 1.0 - por;
    double n = 1.0 - sf;
    s->por[i] = n;
    s->rho[i] = n * s->rho0[i];
    double n3 = (n * n) * n;
    double fch = 1.750 / sqrt(150 * n3);
    double sp = sqrt((s->u[i] * s->u[i]) + (s->v[i] * s->v[i]));
    double a = -((((o_vis(s->rho0[i]) / s->rho0[i]) * lamk) * ((1.0 - n) * (1.0 - n))) /
                 (n3 * (d50 * d50)));
    double b = (((fch * sp) * sqrt(lamk)) * (1.0 - n)) / (sqrt(MUT(6) ? n * n : n3) * d50);
    s->resist[i] = a - b;
  }
}

/* ------------------------------------------------------------------ */
/* eddyvis, turbulence.f90:12-110                                       */
/* ------------------------------------------------------------------ */
void or_eddyvis(ostate *s) {
  const double dl = s->P.dl;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] != ISPH_TID_FLUID) { s->nut[i] = 0.0; s->tke[i] = 0.0; continue; }
    const double *cf = &s->coff[4 * i];
    double t1 = 0.0, t2 = 0.0, t3 = 0.0;
    for (int m = 0; m < s->ncnt[i]; m++) {
      int32_t j = s->nbr[i * KCAP + m];
      double z1 = s->klt1[i * KCAP + m], z2 = s->klt2[i * KCAP + m];
      double gx = cf[0] * z1 + cf[1] * z2, gy = cf[2] * z1 + cf[3] * z2;
      double du = (s->u[j] / s->por[j]) - (s->u[i] / s->por[i]);
      double dv = (s->v[j] / s->por[j]) - (s->v[i] / s->por[i]);
      t1 = t1 + ((s->m[j] * du) * gx) / s->rho[j];
      t2 = t2 + ((s->m[j] * dv) * gy) / s->rho[j];
      t3 = t3 + ((0.50 * s->m[j]) * ((du * gy) + (dv * gx))) / s->rho[j];
    }
    double t4 = sqrt((2 * ((t1 * t1) + (t2 * t2))) + ((2 * t3) * (2 * t3)));
    s->nut[i] = ((0.150 * dl) * (0.150 * dl)) * t4;
    s->tke[i] = (0.080 / 1.0) * ((dl * t4) * (dl * t4));
  }
}

/* ------------------------------------------------------------------ */
/* int_vel, integrator.f90:459-703                                      */
/* ------------------------------------------------------------------ */
void or_int_vel(ostate *s) {
  const double dl = s->P.dl, dt = s->dt, h = s->P.h, lam = s->P.lam, g = s->P.g;
  const double eps6 = (double)1e-6f;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (!(s->flags[i] & ISPH_F_FREE)) continue;
    const double *cf = &s->coff[4 * i];
    double t1 = 0.0, t2 = 0.0;
    s->varts[i] = 0.0;
    for (int m = 0; m < s->ncnt[i]; m++) {
      int32_t j = s->nbr[i * KCAP + m];
      double z1 = s->klt1[i * KCAP + m], z2 = s->klt2[i * KCAP + m];
      double gx = cf[0] * z1 + cf[1] * z2, gy = cf[2] * z1 + cf[3] * z2;
      t1 = t1 + (((s->u[j] / s->por[j]) - (s->u[i] / s->por[i])) * gx) * (s->m[j] / s->rho[j]);
      t2 = t2 + (((s->v[j] / s->por[j]) - (s->v[i] / s->por[i])) * gy) * (s->m[j] / s->rho[j]);
    }
    s->varts[i] = (((1.0 * (dl * dl)) * sqrt((t1 * t1) + (MUT(17) ? 0.0 : t2 * t2))) *
                   sqrt((s->u[i] * s->u[i]) + (s->v[i] * s->v[i]))) /
                  ((s->umax + eps6) * s->por[i]);
  }
  const double wo = o_wo(h);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    int buf = (s->flags[i] & ISPH_F_BUFFER) != 0;
    if (s->tid[i] == ISPH_TID_FLUID && !buf) {
      double us = 0.0, vs = 0.0;
      const double *cf = &s->coff[4 * i];
      const double ni = s->por[i];
      const int freei = (s->flags[i] & ISPH_F_FREE) != 0;
      for (int m = 0; m < s->ncnt[i]; m++) {
        int32_t j = s->nbr[i * KCAP + m];
        double z1 = s->klt1[i * KCAP + m], z2 = s->klt2[i * KCAP + m];
        double r = s->nr[i * KCAP + m];
        double gx = cf[0] * z1 + cf[1] * z2, gy = cf[2] * z1 + cf[3] * z2;
        double cmu;
        if (s->tid[j] == ISPH_TID_FLUID) {
          double vi = o_vis(s->rho0[i]), vj = o_vis(s->rho0[j]);
          cmu = (((2 * vi) * vj) * ((1.0 / s->rho0[i]) * (1.0 / s->rho0[j]))) /
                ((vi / s->rho0[i]) + (vj / s->rho0[j]));
        } else {
          cmu = o_vis(s->rho0[i]) * (1.0 / s->rho0[i]);
        }
        double vart;
        int freej = (s->flags[j] & ISPH_F_FREE) != 0;
        if (freei) vart = s->varts[i];
        else if (!freei && freej && !buf) vart = (o_wab(r, h) * s->varts[j]) / wo;
        else vart = 0.0;
        double pvol = s->m[j] / s->rho[j];
        double pd = 1.0 / ((r * r) + lam);
        double dxij = s->x[i] - s->x[j], dyij = s->y[i] - s->y[j];
        double a1 = (((((MUT(7) ? 1 : 2) * pvol) * (s->u[i] - s->u[j])) * dxij) * gx) * pd;
        double a2 = (((((MUT(7) ? 1 : 2) * pvol) * (s->u[i] - s->u[j])) * dyij) * gy) * pd;
        us = us + (a1 + a2) * (cmu + vart);
        if (s->tid[j] > 2) {
          if ((s->nut[i] < eps6) && (s->nut[j] < eps6))
            us = us + ((a1 + a2) * (s->nut[i] + s->nut[j])) * 0.50;
          else
            us = us + (a1 + a2) * ((((MUT(19) ? 1 : 2) * s->nut[i]) * s->nut[j]) / (s->nut[i] + s->nut[j]));
        }
        a1 = ((((2 * pvol) * (s->v[i] - s->v[j])) * dxij) * gx) * pd;
        a2 = ((((2 * pvol) * (s->v[i] - s->v[j])) * dyij) * gy) * pd;
        vs = vs + (a1 + a2) * (cmu + vart);
        if (s->tid[j] > 2) {
          if ((s->nut[i] < eps6) && (s->nut[j] < eps6))
            vs = vs + ((a1 + a2) * (s->nut[i] + s->nut[j])) * 0.50;
          else
            vs = vs + (a1 + a2) * (((2 * s->nut[i]) * s->nut[j]) / (s->nut[i] + s->nut[j]));
        }
        double tk = (pvol * (s->tke[j] - s->tke[i])) * gx;
        us = MUT(8) ? us + ((tk * ni) * 2.0) / 3 : us - ((tk * ni) * 2.0) / 3;
        tk = (pvol * (s->tke[j] - s->tke[i])) * gy;
        vs = vs - ((tk * ni) * 2.0) / 3;
      }
      double den = MUT(9) ? 1.0 / (1.0 + ((dt * s->resist[i]) * ni)) : 1.0 - ((dt * s->resist[i]) * ni);
      us = (s->u[i] + (dt * us)) / den;
      vs = (s->v[i] + (dt * vs)) / den;
      vs = vs + ((g * ni) * dt);
      s->us[i] = us;
      s->vs[i] = vs;
    }
    if (buf) {
      s->us[i] = s->u[i];
      s->vs[i] = (g * s->por[i]) * dt;
    }
  }
}

/* ------------------------------------------------------------------ */
/* ppesolve assembly, isph.f90:115-456                                   */
/* ------------------------------------------------------------------ */
int or_ppe_assemble(ostate *s) {
  const double lam = s->P.lam, dt = s->dt, rhoref = s->P.rho, rel = s->P.rel_den;
  int bad = 0;
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : bad, err)
  for (int64_t i = 0; i < s->n; i++) {
    double val[RCAP];
    int32_t col[RCAP];
    uint8_t wr[RCAP];
    int num = s->ncnt[i];
    if (num < 0 || num + 1 > RCAP) { err = 1; continue; }
    for (int k = 0; k < RCAP; k++) { val[k] = 0.0; col[k] = (int32_t)i; wr[k] = 0; }
    val[0] = 1.0;
    int dia = 0, ins = 0;
    double diag = 0.0, rhs = 0.0;
    const double ni = s->por[i], rhoi = s->rho[i];
    const double rho_i = rhoi * (1.0 / ni);
    const int tidi = s->tid[i];
    const int freei = (s->flags[i] & ISPH_F_FREE) != 0;
    const double *cf = &s->coff[4 * i];
    for (int m = 0; m < num; m++) {
      int32_t j = s->nbr[i * KCAP + m];
      if (i > j) { dia = m + 1; ins = m; } else { ins = m + 1; }
      double lamp = (freei && s->tid[j] == ISPH_TID_FLUID) ? (1.0 + s->rho0[j] / (rhoref * rel)) : 1.0;
      double nj = s->por[j], rhoj = s->rho[j];
      double rho_j = rhoj * (1.0 / nj);
      double pvol = s->m[j] * (1.0 / rhoj);
      double d = o_dist(s->x[j], s->y[j], s->x[i], s->y[i]);
      double pd = 1.0 / ((d * d) + lam);
      double rho_ij = (((2 * rhoj) * rhoi) * (1.0 / (nj * ni))) *
                      (1.0 / ((rhoj * (1.0 / nj)) + (rhoi * (1.0 / ni))));
      double z1 = s->klt1[i * KCAP + m], z2 = s->klt2[i * KCAP + m];
      double dx = s->x[i] - s->x[j], dy = s->y[i] - s->y[j];
      double t1, t2;
      if (tidi == ISPH_TID_FLUID) {
        double gx = cf[0] * z1 + cf[1] * z2, gy = cf[2] * z1 + cf[3] * z2;
        t1 = (((2 * pvol) * gx) * dx) * pd;
        t2 = (((2 * pvol) * gy) * dy) * pd;
        val[ins] = -(t1 + t2); col[ins] = j; wr[ins] = 1;
        diag = diag + (t1 + t2) * lamp;
        if (s->tid[j] == ISPH_TID_FLUID) {
          t1 = MUT(10) ? (pvol * (((s->us[i] * rho_ij) / ni) - ((s->us[j] * rho_ij) / nj))) * gx
                       : (pvol * (((s->us[j] * rho_ij) / nj) - ((s->us[i] * rho_ij) / ni))) * (MUT(11) ? gy : gx);
          t2 = (pvol * (((s->vs[j] * rho_ij) / nj) - ((s->vs[i] * rho_ij) / ni))) * gy;
        } else {
          t1 = (pvol * (((s->us[j] * rho_i) / ni) - ((s->us[i] * rho_i) / ni))) * gx;
          t2 = (pvol * (((s->vs[j] * rho_i) / ni) - ((s->vs[i] * rho_i) / ni))) * gy;
        }
        rhs = rhs + (t1 + t2) / dt;
      } else if (tidi == ISPH_TID_GHOST) {
        t1 = (((2 * pvol) * z1) * dx) * pd;
        t2 = (((2 * pvol) * z2) * dy) * pd;
        if (s->tid[j] <= 2 || MUT(18)) {
          val[ins] = -(t1 + t2); col[ins] = j; wr[ins] = 1;
          diag = diag + (t1 + t2);
        }
      } else if (tidi == ISPH_TID_FIXED) {
        if (s->tid[j] == ISPH_TID_FLUID) {
          const double *cj = &s->coff[4 * j];
          double gxj = cj[0] * z1 + cj[1] * z2, gyj = cj[2] * z1 + cj[3] * z2;
          t1 = ((((2 * s->m[i]) * gxj) * dx) * pd) / rhoi;
          t2 = ((((2 * s->m[i]) * gyj) * dy) * pd) / rhoi;
          val[ins] = -(t1 + t2); col[ins] = j; wr[ins] = 1;
          diag = diag + (t1 + t2);
          t1 = ((s->m[i] * (((s->us[j] * rho_j) / nj) - ((s->us[i] * rho_j) / ni))) * gxj) / rhoi;
          t2 = ((s->m[i] * (((s->vs[j] * rho_j) / nj) - ((s->vs[i] * rho_j) / ni))) * gyj) / rhoi;
          rhs = rhs + (t1 + t2) / dt;
        } else if (s->tid[j] <= 2) {
          t1 = (((2 * pvol) * z1) * dx) * pd;
          t2 = (((2 * pvol) * z2) * dy) * pd;
          val[ins] = -(t1 + t2); col[ins] = j; wr[ins] = 1;
          diag = diag + (t1 + t2);
        }
      }
    }
    /* no neighbours: the identity row (p = 0), the val(dia_pt) = 1 that
     * isph.f90:425-427 leaves commented out; the live code divides 0 by 0 */
    if (num == 0) diag = 1.0;
    val[dia] = diag; col[dia] = (int32_t)i; wr[dia] = 1;
    double t1 = val[dia];
    rhs = rhs / t1;
    for (int k = 0; k < RCAP; k++) val[k] = val[k] / t1;
    if (!(isfinite(t1) && t1 != 0.0)) bad = 1;
    /* canonical order: ascending columns; the row's diagonal precedes the
     * slot-1 value left from the row initialisation (isph.f90:127). */
    int L = 0;
    int32_t *oc = &s->acol[i * RCAP];
    double *ov = &s->aval[i * RCAP];
    for (int k = 0; k < dia; k++) if (wr[k]) { oc[L] = col[k]; ov[L] = val[k]; L++; }
    oc[L] = col[dia]; ov[L] = val[dia]; L++;
    if (dia > 0 && !wr[0]) { oc[L] = col[0]; ov[L] = val[0]; L++; }
    for (int k = dia + 1; k <= num; k++) if (wr[k]) { oc[L] = col[k]; ov[L] = val[k]; L++; }
    s->alen[i] = L;
    s->b[i] = rhs;
  }
  s->assembled = 1;
  if (err) return ISPH_ECAP;
  return bad ? ISPH_ENONFINITE : 0;
}

/* ------------------------------------------------------------------ */
/* solve: spmv2 (solver.f90:84-127), BiCGStab (solver.f90:129-254)     */
/* ------------------------------------------------------------------ */
/* Blocked tree reduction shared with the device: blocks of RED_BLOCK
 * consecutive rows are summed by a halving tree (v[t] += v[t+s] for
 * s = 128..1); block partials are then summed by 256 sequential lanes
 * (lane t adds partials t, t+256, ...) followed by the same tree. */
static double tree256(double *v) {
  for (int st = RED_BLOCK / 2; st >= 1; st >>= 1)
    for (int t = 0; t < st; t++) v[t] = v[t] + v[t + st];
  return v[0];
}

static double reduce_prod(const ostate *s, int64_t n, const double *prod) {
  (void)s;
  int64_t nb = (n + RED_BLOCK - 1) / RED_BLOCK;
  double *part = malloc(sizeof(double) * (nb > 0 ? nb : 1));
#pragma omp parallel for schedule(static)
  for (int64_t bk = 0; bk < nb; bk++) {
    double v[RED_BLOCK];
    for (int t = 0; t < RED_BLOCK; t++) {
      int64_t r = bk * RED_BLOCK + t;
      v[t] = (r < n) ? prod[r] : 0.0;
    }
    part[bk] = tree256(v);
  }
  double v[RED_BLOCK];
  for (int t = 0; t < RED_BLOCK; t++) {
    double acc = 0.0;
    for (int64_t k = t; k < nb; k += RED_BLOCK) acc = acc + part[k];
    v[t] = acc;
  }
  free(part);
  return tree256(v);
}

static double dot(const ostate *s, int64_t n, const double *a, const double *b, double *tmp) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) tmp[i] = a[i] * b[i];
  return reduce_prod(s, n, tmp);
}

static void spmv(const ostate *s, const double *xv, double *yv) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    double acc = 0.0;
    const int32_t *c = &s->acol[i * RCAP];
    const double *a = &s->aval[i * RCAP];
    for (int k = 0; k < s->alen[i]; k++) acc = acc + a[k] * xv[c[k]];
    yv[i] = acc;
  }
}

/* FP64 BiCGStab in the reference's update order (solver.f90:182-213). */
/* BiCGStab in the update order of solver.f90:182-213, stopping on the
 * relative residual.  On a breakdown (rho, omega or alpha zero or not
 * finite) below the iteration cap it restarts from the true residual
 * (r = b - A x, rh = p = r), at most max_outer times; iterations keep
 * counting across restarts and st->outer counts the restarts.  (The
 * reference's bicgstab has neither the residual stop nor the restart.) */
static int bicgstab64(ostate *s, double *xv, const double *bv, double tol, int maxit,
                      int max_outer, isph_solve_stats *st, double bnorm) {
  const int64_t n = s->n;
  double *r = s->w1, *rh = s->w2, *p = s->w3, *t1 = s->w4, *sv = s->w5, *t2 = s->w6, *tmp = s->w7;
  st->iterations = 0;
  st->breakdown = 0;
  st->converged = 0;
  st->outer = 0;
  int it = 0;
  for (int restart = 0;; restart++) {
    spmv(s, xv, t1);
    for (int64_t i = 0; i < n; i++) { r[i] = bv[i] - t1[i]; rh[i] = r[i]; p[i] = r[i]; }
    double rho = dot(s, n, rh, r, tmp);
    double rr = dot(s, n, r, r, tmp);
    st->rel_residual = sqrt(rr) / bnorm;
    st->breakdown = 0;
    if (sqrt(rr) <= tol * bnorm) { st->converged = 1; return 0; }
    int brk = 0;
    while (it < maxit) {
      it++;
      spmv(s, p, t1);
      double alpha = rho / dot(s, n, t1, rh, tmp);
      for (int64_t i = 0; i < n; i++) sv[i] = r[i] - alpha * t1[i];
      spmv(s, sv, t2);
      double omega = dot(s, n, t2, sv, tmp) / dot(s, n, t2, t2, tmp);
      for (int64_t i = 0; i < n; i++) {
        xv[i] = (xv[i] + alpha * p[i]) + omega * sv[i];
        r[i] = sv[i] - omega * t2[i];
      }
      double rhon = dot(s, n, r, rh, tmp);
      rr = dot(s, n, r, r, tmp);
      st->iterations = it;
      st->rel_residual = sqrt(rr) / bnorm;
      if (sqrt(rr) <= tol * bnorm) { st->converged = 1; return 0; }
      if (!(isfinite(alpha) && isfinite(omega)) || omega == 0.0 || rhon == 0.0 || !isfinite(rhon)) {
        brk = 1;
        break;
      }
      double beta = (rhon / rho) * (alpha / omega);
      for (int64_t i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * t1[i]);
      rho = rhon;
    }
    if (!brk) return 0;            /* iteration cap */
    st->breakdown = 1;
    if (it >= maxit || restart >= max_outer) return 0;
    st->outer = restart + 1;
  }
}

int or_ppe_solve(ostate *s, isph_solve_stats *out) {
  if (!s->assembled) return ISPH_ESTATE;
  const int64_t n = s->n;
  isph_solve_stats st;
  memset(&st, 0, sizeof st);
  double bn2 = dot(s, n, s->b, s->b, s->w8);
  double bnorm = sqrt(bn2);
  st.bnorm = bnorm;
  for (int64_t i = 0; i < n; i++) st.nnz += s->alen[i];
  if (bnorm == 0.0 || !isfinite(bnorm)) {
    if (bnorm == 0.0) {
      for (int64_t i = 0; i < n; i++) s->sol[i] = 0.0;
      st.converged = 1;
    }
  } else {
    /* The oracle is the FP64 reference for both solve forms: the precision
     * form of the GPU path is checked against it by tolerance. */
    bicgstab64(s, s->sol, s->b, s->P.tol, s->P.max_iter, s->P.max_outer, &st, bnorm);
  }
  /* pressure scatter, isph.f90:468-486 */
  for (int64_t i = 0; i < n; i++) s->p[i] = s->sol[i];
  s->last = st;
  if (out) *out = st;
  if (!isfinite(st.rel_residual) || !isfinite(bnorm)) return ISPH_ENONFINITE;
  return st.converged ? 0 : ISPH_ENOCONV;
}

/* ------------------------------------------------------------------ */
/* comp_vel, comp_pos (integrator.f90:705-883)                          */
/* ------------------------------------------------------------------ */
/* per-cell maxima of |vel| and |nu_t| (integrator.f90:831-841): a cell with
 * particles takes the maxima over its slots, an empty cell keeps its old
 * values.  vel counts for the non-buffer fluid particles comp_vel updated
 * (walls: the never-written slot value 0; a buffer particle: 0, where the
 * reference reads its slot's stale vel).  NaN counts as 0.  init: the
 * velocities count as 0 and empty cells are cleared. */
static double nan0(double a) { return (a == a) ? fabs(a) : 0.0; }
static void cell_maxima(ostate *s, int init) {
  for (int32_t k = 0; k < s->ncell; k++) {
    int32_t i0 = s->cstart[k], i1 = s->cstart[k + 1];
    if (i1 == i0) {
      if (init) { s->cmaxvel[k] = 0.0; s->cmaxeddy[k] = 0.0; }
      continue;
    }
    double mv = 0.0, mn = 0.0;
    for (int32_t i = i0; i < i1; i++) {
      if (!init && s->tid[i] == ISPH_TID_FLUID && !(s->flags[i] & ISPH_F_BUFFER)) {
        double a = nan0(s->vel[i]);
        if (a > mv) mv = a;
      }
      double b = nan0(s->nut[i]);
      if (b > mn) mn = b;
    }
    s->cmaxvel[k] = mv;
    s->cmaxeddy[k] = mn;
  }
}

void or_comp_vel(ostate *s) {
  const double dt = s->dt;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    int buf = (s->flags[i] & ISPH_F_BUFFER) != 0;
    if (s->tid[i] == ISPH_TID_FLUID && !buf) {
      double u = 0.0, v = 0.0;
      const double *cf = &s->coff[4 * i];
      const double ni = s->por[i], rhoi = s->rho[i], pi = s->p[i];
      for (int m = 0; m < s->ncnt[i]; m++) {
        int32_t j = s->nbr[i * KCAP + m];
        double z1 = s->klt1[i * KCAP + m], z2 = s->klt2[i * KCAP + m];
        double gx = cf[0] * z1 + cf[1] * z2, gy = cf[2] * z1 + cf[3] * z2;
        double nj = s->por[j], rhoj = s->rho[j], pj = s->p[j];
        double V = s->m[j] / rhoj, t1;
        if (s->tid[j] == ISPH_TID_FLUID) {
          double a = ((pj * rhoi) / ni) + ((pi * rhoj) / nj);
          double den = rhoi * ((rhoi / ni) + (rhoj / nj));
          const double nn = MUT(2) ? ni : (ni * ni);
          t1 = ((MUT(1) ? 1 : 2) * (((V * nn) * a) * gx)) / den;
          u = u - t1 * dt;
          t1 = ((MUT(1) ? 1 : 2) * (((V * nn) * a) * gy)) / den;
          v = v - t1 * dt;
        } else {
          t1 = (((V * (ni * ni)) * (pj + pi)) * gx) / rhoi;
          u = u - t1 * dt;
          t1 = (((V * (ni * ni)) * (pj + pi)) * gy) / rhoi;
          v = v - t1 * dt;
        }
      }
      u = s->us[i] + u;
      v = s->vs[i] + v;
      s->u[i] = u;
      s->v[i] = v;
      s->vel[i] = sqrt((u * u) + (v * v)) / ni;
    }
    if (buf) { s->u[i] = s->us[i]; s->v[i] = 0.0; }
  }
  cell_maxima(s, 0);
}

void or_comp_pos(ostate *s) {
  const double dt = s->dt;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] != ISPH_TID_FLUID) continue;
    s->x[i] = ((s->x[i] + (MUT(15) ? s->x[i] : s->xs[i])) + ((dt * s->u[i]) / s->por[i])) / 2.0;
    s->y[i] = ((s->y[i] + s->ys[i]) + ((dt * s->v[i]) / s->por[i])) / 2.0;
    if ((s->flags[i] & ISPH_F_BUFFER) && (s->x[i] > s->P.buf_xmin) && (s->x[i] < s->P.buf_xmax))
      s->flags[i] &= (uint8_t)~ISPH_F_BUFFER;
  }
}

/* freesurf, isph.f90:49-86: vicinity flags of the neighbours of surface
 * particles (sequential: the flag writes go to other particles). */
void or_freesurf(ostate *s) {
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] == ISPH_TID_FLUID && !(s->flags[i] & ISPH_F_BUFFER) &&
        s->gradvx[i] <= s->P.fs_thresh) {
      for (int m = 0; m < s->ncnt[i]; m++)
        s->flags[MUT(16) ? i : s->nbr[i * KCAP + m]] |= ISPH_F_VICINITY;
    }
  }
}

/* ------------------------------------------------------------------ */
/* opt2_shift, part_shift.f90:12-191                                    */
/* ------------------------------------------------------------------ */
void or_opt2_shift(ostate *s) {
  const double h = s->P.h, dl = s->P.dl, csh = s->P.csh, cap = s->P.maxshift * dl;
  const double wdl = o_wab(dl, h);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (!(s->tid[i] == ISPH_TID_FLUID && !(s->flags[i] & ISPH_F_BUFFER))) continue;
    int cnt = s->ncnt[i];
    if (cnt == 0) {
      /* no neighbour: no shift (the reference stores its zeroed local
       * shifts only for count > 0, part_shift.f90:32-35, 38-161, and would
       * add the slot's saved position and intermediate velocity) */
      s->xs[i] = 0.0; s->ys[i] = 0.0; s->us[i] = 0.0; s->vs[i] = 0.0;
      continue;
    }
    const double *cf = &s->coff[4 * i];
    /* the surface normal of part_shift.f90:50-65 is overwritten by zero at
     * part_shift.f90:67-69 before any use; pint is zero as well */
    double normx = 0.0, normy = 0.0;
    int pint = 0;
    double frac = (s->flags[i] & ISPH_F_FREE) ? 0.01 : 1.0;
    double xsa = 0.0, ysa = 0.0;
    for (int m = 0; m < cnt; m++) {
      int32_t j = s->nbr[i * KCAP + m];
      double z1 = s->klt1[i * KCAP + m], z2 = s->klt2[i * KCAP + m];
      double wq = o_wab(s->nr[i * KCAP + m], h) / wdl;
      double f = 1.0 + (0.20 * pint) * ((wq * wq) * (wq * wq));
      const double dxs = (((((csh * (h * h)) * frac) * (((1 - normx * normx) * z1) - ((normx * normy) * z2))) *
                           (s->m[j] / s->rho[j])) * f);
      xsa = MUT(3) ? xsa + dxs : xsa - dxs;
    }
    for (int m = 0; m < cnt; m++) {
      int32_t j = s->nbr[i * KCAP + m];
      double z1 = s->klt1[i * KCAP + m], z2 = s->klt2[i * KCAP + m];
      double wq = o_wab(s->nr[i * KCAP + m], h) / wdl;
      double f = 1.0 + (0.20 * pint) * ((wq * wq) * (wq * wq));
      ysa = ysa - (((((csh * (h * h)) * frac) * (((-(normx * normy)) * z1) + ((1 - normy * normy) * z2))) *
                   (s->m[j] / s->rho[j])) * f);
    }
    s->xs[i] = (xsa == 0.0) ? 0.0 : (o_min(fabs(xsa), cap) * xsa) / fabs(xsa);
    s->ys[i] = (ysa == 0.0) ? 0.0 : (o_min(fabs(ysa), cap) * ysa) / fabs(ysa);
    double t1 = 0.0, t2 = 0.0;
    for (int m = 0; m < cnt; m++) {
      int32_t j = s->nbr[i * KCAP + m];
      double z1 = s->klt1[i * KCAP + m], z2 = s->klt2[i * KCAP + m];
      double gx = cf[0] * z1 + cf[1] * z2, gy = cf[2] * z1 + cf[3] * z2;
      double V = s->m[j] / s->rho[j];
      t1 = t1 + (gx * (s->u[j] - s->u[i])) * V;
      t2 = t2 + (gy * (s->u[j] - s->u[i])) * V;
    }
    s->us[i] = MUT(4) ? t2 * xsa + t1 * ysa : t1 * xsa + t2 * ysa;
    t1 = 0.0;
    t2 = 0.0;
    for (int m = 0; m < cnt; m++) {
      int32_t j = s->nbr[i * KCAP + m];
      double z1 = s->klt1[i * KCAP + m], z2 = s->klt2[i * KCAP + m];
      double gx = cf[0] * z1 + cf[1] * z2, gy = cf[2] * z1 + cf[3] * z2;
      double V = s->m[j] / s->rho[j];
      t1 = t1 + (gx * (s->v[j] - s->v[i])) * V;
      t2 = t2 + (gy * (s->v[j] - s->v[i])) * V;
    }
    s->vs[i] = t1 * xsa + t2 * ysa;
  }
  /* shifting hydrodynamic values, part_shift.f90:171-189 */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] == ISPH_TID_FLUID && !(s->flags[i] & ISPH_F_BUFFER)) {
      s->x[i] = s->x[i] + s->xs[i];
      s->y[i] = s->y[i] + s->ys[i];
      s->u[i] = s->u[i] + s->us[i];
      s->v[i] = s->v[i] + s->vs[i];
    }
  }
}

/* densityupdate, scalar.f90:399-426 */
void or_densityupdate(ostate *s) {
  const double rhomax = s->P.rho * s->P.rel_den, rhomin = s->P.rhomin;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] != ISPH_TID_FLUID) continue;
    s->rho0[i] = (0.5 * (rhomax + rhomin)) + (s->c[i] * (rhomax - rhomin));
    s->m[i] = s->rho0[i] * s->vol0[i];
    s->rho[i] = s->rho0[i] * s->por[i];
  }
}

/* timestep, integrator.f90:885-953 */
void or_timestep(ostate *s) {
  const isph_params *P = &s->P;
  /* maxima over every cell, emptied ones included (integrator.f90:914-924) */
  s->umax = 0.0;
  s->numax = 0.0;
  for (int32_t k = 0; k < s->ncell; k++) {
    if (s->cmaxvel[k] > s->umax) s->umax = s->cmaxvel[k];
    if (s->cmaxeddy[k] > s->numax) s->numax = s->cmaxeddy[k];
  }
  const double dl = P->dl, sig1 = P->sig1;
  if (!P->fixed_dt) {
    double t = s->t, dtn;
    if (t <= 1.0) {
      dtn = o_min(o_min(sig1 * dl / s->umax, (sig1 * (dl * dl)) / ((P->mu / P->rho) + s->numax)), 0.0010);
    } else if ((t > 1.0) && (t <= 2.50)) {
      float a = (float)(sig1 * dl / s->umax), b = (float)(((sig1 * (dl * dl)) * P->rho) / P->mu);
      float mn = a; if (b < mn) mn = b; if (0.002f < mn) mn = 0.002f;
      dtn = (double)mn;
    } else if ((t > 3.0) && (t <= 4.0)) {
      float a = (float)(sig1 * dl / s->umax), b = (float)(((sig1 * (dl * dl)) * P->rho) / P->mu);
      float mn = a; if (b < mn) mn = b; if (0.004f < mn) mn = 0.004f;
      dtn = (double)mn;
    } else {
      dtn = o_min(o_min(sig1 * dl / s->umax, (sig1 * (dl * dl)) / ((P->mu / P->rho) + s->numax)), 0.0050);
    }
    s->dt = dtn;
  }
  if (!isfinite(s->numax) || isnan(s->numax)) s->numax = 0.0;
  double dtsol = (sig1 * (dl * dl)) / ((double)1e-9f + s->numax / P->tschmidt);
  int solsteps = 1;
  if (dtsol < s->dt) {
    solsteps = (int)ceil(s->dt / dtsol);
    dtsol = s->dt / solsteps;
  }
  s->dtsol = (dtsol < s->dt) ? dtsol : s->dt;
  s->solsteps = solsteps;
}

/* ------------------------------------------------------------------ */
/* scalart_anisotropic / scalarupdate, scalar.f90:242-397               */
/* ------------------------------------------------------------------ */
void or_scalart(ostate *s) {
  const double al = 0.0025, at = 0.00025, Dm = (double)1e-9f, thr = 0.7;
  const double h = s->P.h, lam = s->P.lam, dl = s->P.dl, dt = s->dt, tsc = s->P.tschmidt;
  const double eps6 = (double)1e-6f;
  const double capv = (0.12 * (dl * dl)) / dt;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] != ISPH_TID_FLUID) continue;
    double cd = 0.0;
    const double c1 = s->c[i], ni = s->por[i];
    for (int m = 0; m < s->ncnt[i]; m++) {
      int32_t j = s->nbr[i * KCAP + m];
      double r = s->nr[i * KCAP + m];
      double vi = sqrt((s->u[i] * s->u[i]) + (s->v[i] * s->v[i])) / ni;
      if (s->tid[j] != ISPH_TID_FLUID) continue;
      double c2 = s->c[j], nj = s->por[j];
      double vj = sqrt((s->u[j] * s->u[j]) + (s->v[j] * s->v[j])) / nj;
      int an = (ni < thr) && (nj < thr);
      double Dxxi, Dxyi, Dyxi, Dyyi, Dxxj, Dxyj, Dyxj, Dyyj, t1x;
      if (an) {
        Dxxi = s->u[i] / ((sqrt((al * vi) + Dm) * vi) * ni);
        Dxyi = s->v[i] / ((sqrt((al * vi) + Dm) * vi) * ni);
        Dyxi = (-s->v[i]) / ((sqrt((at * vi) + Dm) * vi) * ni);
        Dyyi = s->u[i] / ((sqrt((at * vi) + Dm) * vi) * ni);
        Dxxj = s->u[j] / ((sqrt((al * vj) + Dm) * vj) * nj);
        Dxyj = s->v[j] / ((sqrt((al * vj) + Dm) * vj) * nj);
        Dyxj = (-s->v[j]) / ((sqrt((at * vj) + Dm) * vj) * nj);
        Dyyj = s->u[j] / ((sqrt((at * vj) + Dm) * vj) * nj);
        double pd = 1.0 / r;
        double dx = s->x[i] - s->x[j], dy = s->y[i] - s->y[j];
        double A = ((((Dxxi + Dxxj) * dx) * pd) * 0.5) + ((((Dxyi + Dxyj) * dy) * pd) * 0.5);
        double B = ((((Dyxi + Dyxj) * dx) * pd) * 0.5) + ((((Dyyi + Dyyj) * dy) * pd) * 0.5);
        t1x = (A * A) + (B * B);
        t1x = 1.0 / t1x;
        t1x = (t1x <= capv) ? t1x : capv;
      } else {
        Dxxi = Dm + s->nut[i] / tsc;
        Dxxj = Dm + s->nut[j] / tsc;
        if ((s->nut[i] < eps6) && (s->nut[j] < eps6)) t1x = (Dxxi + Dxxj) * (MUT(12) ? 1.0 : 0.50);
        else t1x = ((2 * Dxxi) * Dxxj) / (Dxxi + Dxxj);
      }
      double t1y = t1x, t2 = 1.0;
      double t3 = (s->vol0[i] * (ni + nj)) / (ni * nj);
      if (MUT(13)) t3 = -t3;
      double dx = s->x[i] - s->x[j], dy = s->y[i] - s->y[j];
      double hwx = o_wabd(dx, r, h), hwy = o_wabd(dy, r, h);
      double den = (r * r) + lam;
      cd = (cd + (((((((t1x * t2) * t3) * (c1 - c2)) * dx) * hwx) / den))) +
           (((((((t1y * t2) * t3) * (c1 - c2)) * dy) * hwy) / den));
    }
    s->cdiff[i] = cd;
  }
}

void or_scalarupdate(ostate *s, double stime) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] != ISPH_TID_FLUID) continue;
    s->c[i] = s->c[i] + s->cdiff[i] * stime;
    if ((s->flags[i] & ISPH_F_BUFFER) && (s->x[i] < 1.6)) s->c[i] = 0.5;
  }
}

/* ------------------------------------------------------------------ */
/* drivers                                                              */
/* ------------------------------------------------------------------ */
/* State at the start of the production loop: lists, porosity, nut/tke. */
int or_init_state(ostate *s) {
  int rc = or_neighbour(s);
  if (rc) return rc;
  or_effpor(s);
  or_eddyvis(s);
  cell_maxima(s, 1);
  or_timestep(s);
  return 0;
}

/* One production-loop iteration, main.f90:199-261. */
/* ------------------------------------------------------------------ */
/* inflow/outflow buffers (internal.f90:447-586, main.f90:217-235)        */
/* ------------------------------------------------------------------ */
/* Zone 1 or 2.  The template particles (x, y, vx, vy, pressure) come in the
 * reference's loop order (buffer(k,i), k fastest); their cells are the zone's
 * entry cells.  An inserted particle takes the template's position, velocity
 * and pressure and the zone's mass, density (also the reference density),
 * volume and concentration (insert_buffer1: fmass, rho, fmass/rho, -0.5;
 * insert_buffer2: fmass*relden, rho*relden, fmass/rho, +0.5).  rv: zone 1
 * removes buffer particles with x >= rv (remove_buffer1, rv_buf_r), zone 2
 * those with x <= rv (remove_buffer2, rv_buf_l).  interval (ins_1, ins_2)
 * and time_shift drive the insertion clock; interval 0 leaves it to
 * or_buffer_update. */
int or_set_buffer(ostate *s, int zone, int64_t nt, const double *x, const double *y,
                  const double *u, const double *v, const double *p, double mass, double rho,
                  double vol0, double c, double rv, double interval, double time_shift) {
  if (zone < 1 || zone > 2 || nt < 0) return ISPH_EARG;
  ozone *Z = &s->zone[zone - 1];
  free(Z->tx); free(Z->ty); free(Z->tu); free(Z->tv); free(Z->tp); free(Z->tcell);
  free(Z->entry);
  size_t m = (size_t)(nt > 0 ? nt : 1);
  Z->tx = malloc(8 * m); Z->ty = malloc(8 * m); Z->tu = malloc(8 * m);
  Z->tv = malloc(8 * m); Z->tp = malloc(8 * m); Z->tcell = malloc(4 * m);
  Z->entry = calloc((size_t)s->ncell, 1);
  for (int64_t t = 0; t < nt; t++) {
    int ci, cj;
    int k = locate(s, x[t], y[t], &ci, &cj);
    if (k < 0) return ISPH_EARG;
    Z->tx[t] = x[t]; Z->ty[t] = y[t]; Z->tu[t] = u[t]; Z->tv[t] = v[t]; Z->tp[t] = p[t];
    Z->tcell[t] = k;
    Z->entry[k] = 1;
  }
  Z->nt = nt;
  Z->mass = mass; Z->rho = rho; Z->vol0 = vol0; Z->c = c; Z->rv = rv;
  Z->interval = interval; Z->time_shift = time_shift;
  Z->iter = 1;
  Z->set = 1;
  return 0;
}

/* remove_buffer then insert_buffer of one zone: in every entry cell the
 * kept particles stay in order, buffer particles past rv are dropped, and
 * the zone's template particles of that cell are appended in template
 * order; other cells are untouched.  Porosity moves with its particle
 * (inserted particles get 1); the other derived properties are recomputed
 * before they are read and are poisoned here as after a rebin.  Solution
 * rows beyond the old particle count start from 0. */
int or_buffer_update(ostate *s, int zone) {
  if (zone < 1 || zone > 2 || !s->zone[zone - 1].set) return ISPH_EARG;
  const ozone *Z = &s->zone[zone - 1];
  if (s->n + Z->nt > s->cap) return ISPH_ECAP;
  int64_t *src = malloc(sizeof(int64_t) * (size_t)(s->n + Z->nt + 1));
  int64_t nn = 0;
  for (int k = 0; k < s->ncell; k++) {
    for (int64_t i = s->cstart[k]; i < s->cstart[k + 1]; i++) {
      int drop = 0;
      if (Z->entry[k] && (s->flags[i] & ISPH_F_BUFFER))
        drop = (zone == 1) ? (s->x[i] >= Z->rv) : (s->x[i] <= Z->rv);
      if (!drop) src[nn++] = i;
    }
    if (Z->entry[k])
      for (int64_t t = 0; t < Z->nt; t++)
        if (Z->tcell[t] == k) src[nn++] = -1 - t;
  }
  /* gather into the scratch vectors, field by field */
  double *dfield[] = {s->x, s->y, s->u, s->v, s->us, s->vs, s->xs, s->ys, s->p, s->m,
                      s->rho, s->rho0, s->vol0, s->c, s->por};
  const int nd = (int)(sizeof(dfield) / sizeof(dfield[0]));
  double *tmp = malloc(sizeof(double) * (size_t)(nn > 0 ? nn : 1));
  for (int f = 0; f < nd; f++) {
    for (int64_t d = 0; d < nn; d++) {
      const int64_t q = src[d];
      if (q >= 0) { tmp[d] = dfield[f][q]; continue; }
      const int64_t t = -1 - q;
      double v;
      switch (f) {
        case 0: v = Z->tx[t]; break;
        case 1: v = Z->ty[t]; break;
        case 2: v = Z->tu[t]; break;
        case 3: v = Z->tv[t]; break;
        case 8: v = Z->tp[t]; break;
        case 9: v = Z->mass; break;
        case 10: case 11: v = Z->rho; break;
        case 12: v = Z->vol0; break;
        case 13: v = Z->c; break;
        case 14: v = 1.0; break;
        default: v = 0.0; break;   /* us, vs, xs, ys */
      }
      tmp[d] = v;
    }
    memcpy(dfield[f], tmp, sizeof(double) * (size_t)nn);
  }
  int32_t *itmp = malloc(sizeof(int32_t) * (size_t)(nn > 0 ? nn : 1));
  uint8_t *btmp = malloc((size_t)(nn > 0 ? nn : 1));
  for (int64_t d = 0; d < nn; d++)
    itmp[d] = src[d] >= 0 ? s->tid[src[d]] : ISPH_TID_FLUID;
  memcpy(s->tid, itmp, sizeof(int32_t) * (size_t)nn);
  for (int64_t d = 0; d < nn; d++)
    itmp[d] = src[d] >= 0 ? s->cell[src[d]] : Z->tcell[-1 - src[d]];
  memcpy(s->cell, itmp, sizeof(int32_t) * (size_t)nn);
  for (int64_t d = 0; d < nn; d++)
    btmp[d] = src[d] >= 0 ? s->flags[src[d]] : (uint8_t)ISPH_F_BUFFER;
  memcpy(s->flags, btmp, (size_t)nn);
  for (int64_t i = s->n; i < nn; i++) s->sol[i] = 0.0;
  /* keep the porosity gathered above through the poisoning */
  double *por = tmp;
  memcpy(por, s->por, sizeof(double) * (size_t)nn);
  s->n = nn;
  rebuild_starts(s);
  poison_props(s);
  memcpy(s->por, por, sizeof(double) * (size_t)nn);
  s->assembled = 0;
  free(src); free(tmp); free(itmp); free(btmp);
  return 0;
}

/* ------------------------------------------------------------------ */
/* WCSPH variant (wcsph.f90:13-154; dead and not compiling in the       */
/* reference, restated from its equations)                               */
/* ------------------------------------------------------------------ */
void or_set_wcsph(ostate *s, double c0, double gam, double delt, double swl, int in_step) {
  s->c0 = c0; s->gam = gam; s->delt = delt; s->swl = swl; s->wcsph = in_step;
}

/* density of the hydrostatic pressure at height y (the term2 expression of
 * calmasscons, wcsph.f90:87-91, with swl = brrealy*distfac*(2*bl-1) + wc) */
static double rho_hyd(const ostate *s, double y) {
  const double rho = s->P.rho;
  return rho * pow(((((s->swl - y) * rho) * fabs(s->P.g)) * s->gam) / (rho * (s->c0 * s->c0)) + 1.0,
                   1.0 / s->gam);
}

/* presstoden, wcsph.f90:13-33: density from pressure, every particle */
void or_presstoden(ostate *s) {
  const double rho = s->P.rho;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++)
    s->rho[i] = rho * pow(((s->p[i] * s->gam) / (rho * (s->c0 * s->c0))) + 1.0, 1.0 / s->gam);
}

/* dentopress, wcsph.f90:35-63: Tait pressure of the fluid, clipped at 0
 * (then the density is reset to rho) */
void or_dentopress(ostate *s) {
  const double rho = s->P.rho;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] <= 2) continue;
    s->p[i] = ((rho * (s->c0 * s->c0)) * (pow(s->rho[i] / rho, s->gam) - 1.0)) / s->gam;
    if (s->p[i] < 0.0) {
      s->p[i] = 0.0;
      s->rho[i] = rho;
    }
  }
}

/* calmasscons, wcsph.f90:65-130: continuity with the delta-SPH diffusion
 * term taken relative to the hydrostatic density,
 *   drho_i/dt = sum_j [ -m_j (u_j - u_i) . gradW~ (rho_i / rho_j)
 *               + delt 2h c0 (rho_j - rho_i - (rhoH_j - rhoH_i)) (x_j - x_i) . gradW~
 *                 (m_j / rho_j) / (r^2 + lam) ],
 * per component in the reference's expression order, gradW~ the corrected
 * gradient with respect to x_i (klt = Wabx(neighbour, particle),
 * search.f90:70-73).  Two departures from the routine as written (it is never
 * called there): it keeps the y-component sum in the variable that also holds
 * the hydrostatic difference, so its y-sum is overwritten at every neighbour
 * (here the y-sum has its own accumulator); and it takes (x_i - x_j), which
 * makes the delta term anti-diffusive, where the delta-SPH scheme
 * (Molteni & Colagrossi 2009) takes (x_j - x_i). */
void or_calmasscons(ostate *s) {
  const double h = s->P.h, lam = s->P.lam, c0 = s->c0, delt = s->delt, dt = s->dt;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] <= 2) continue;
    const double *cf = &s->coff[4 * i];
    const double rhi = rho_hyd(s, s->y[i]);
    double tx = 0.0, ty = 0.0;
    for (int mm = 0; mm < s->ncnt[i]; mm++) {
      const int32_t j = s->nbr[i * KCAP + mm];
      const double r = s->nr[i * KCAP + mm];
      const double z1 = s->klt1[i * KCAP + mm], z2 = s->klt2[i * KCAP + mm];
      const double gx = cf[0] * z1 + cf[1] * z2, gy = cf[2] * z1 + cf[3] * z2;
      const double hyd = rho_hyd(s, s->y[j]) - rhi;
      const double vj = s->m[j] / s->rho[j];
      const double ratio = s->rho[i] / s->rho[j];
      const double dif = (s->rho[j] - s->rho[i]) - hyd;
      const double den = r * r + lam;
      tx = (tx - ((s->m[j] * (s->u[j] - s->u[i])) * gx) * ratio) +
           ((((((delt * 2) * h) * c0) * dif) * (s->x[j] - s->x[i])) * gx) * vj / den;
      ty = (ty - ((s->m[j] * (s->v[j] - s->v[i])) * gy) * ratio) +
           ((((((delt * 2) * h) * c0) * dif) * (s->y[j] - s->y[i])) * gy) * vj / den;
    }
    s->tden[i] = s->rho[i] + (tx + ty) * dt;
  }
}

/* updateden, wcsph.f90:132-152 */
void or_updateden(ostate *s) {
  for (int64_t i = 0; i < s->n; i++)
    if (s->tid[i] > 2) s->rho[i] = s->tden[i];
}

/* ------------------------------------------------------------------ */
/* piston wavemaker (no counterpart in the reference)                   */
/* ------------------------------------------------------------------ */

/* X(t) = (a r) sin(w tau), U = dX/dt with the linear start-up ramp r(tau):
 * the paddle displacement and velocity of a piston wavemaker of stroke 2a
 * (Dean & Dalrymple, Water Wave Mechanics, sec. 6.2). */
int or_wavemaker_motion(double stroke, double period, double ramp, double t0, double t,
                        double *x, double *u) {
  if (!x || !u || !(period > 0.0) || !(stroke >= 0.0) || !(ramp >= 0.0)) return ISPH_EARG;
  double tau = t - t0;
  if (!(tau > 0.0)) {
    *x = 0.0;
    *u = 0.0;
    return 0;
  }
  double a = 0.5 * stroke, w = (2.0 * OR_PI) / period, r = 1.0, rp = 0.0;
  if (ramp > 0.0 && tau < ramp) {
    r = tau / ramp;
    rp = 1.0 / ramp;
  }
  double sn = sin(w * tau), cs = cos(w * tau);
  *x = (a * r) * sn;
  *u = a * (((r * w) * cs) + (rp * sn));
  return 0;
}

static int paddle_to(ostate *s, double t) {
  double X, U;
  if (or_wavemaker_motion(s->wv_stroke, s->wv_period, s->wv_ramp, s->wv_t0, t, &X, &U))
    return ISPH_EARG;
  for (int64_t i = 0; i < s->n; i++) {
    if (s->tid[i] == ISPH_TID_FLUID || !(s->flags[i] & ISPH_F_PADDLE)) continue;
    s->x[i] = s->xs[i] + X;
    s->u[i] = U;
    s->v[i] = 0.0;
    /* the PPE source term reads a wall neighbour's intermediate velocity */
    s->us[i] = U;
    s->vs[i] = 0.0;
  }
  return 0;
}

int or_set_wavemaker(ostate *s, double stroke, double period, double ramp, double t0) {
  double X, U;
  if (or_wavemaker_motion(stroke, period, ramp, t0, s->t, &X, &U)) return ISPH_EARG;
  for (int64_t i = 0; i < s->n; i++)
    if (s->tid[i] != ISPH_TID_FLUID && (s->flags[i] & ISPH_F_PADDLE)) s->xs[i] = s->x[i] - X;
  s->wv_stroke = stroke;
  s->wv_period = period;
  s->wv_ramp = ramp;
  s->wv_t0 = t0;
  s->wave = 1;
  return 0;
}

int or_wavemaker(ostate *s) {
  if (!s->wave) return ISPH_ESTATE;
  return paddle_to(s, s->t);
}

/* the insertion clock of main.f90:217-235, told = t - dt */
static int buffer_clock(ostate *s, double told) {
  for (int z = 0; z < 2; z++) {
    ozone *Z = &s->zone[z];
    if (!Z->set || !(Z->interval > 0.0)) continue;
    if (((told + Z->time_shift) < Z->iter * Z->interval) &&
        ((s->t + Z->time_shift) >= Z->iter * Z->interval)) {
      int rc = or_buffer_update(s, z + 1);
      if (rc) return rc;
      Z->iter = Z->iter + 1;
    }
  }
  return 0;
}

int or_step(ostate *s, int nsteps, isph_solve_stats *st) {
  int rcs = 0;
  for (int k = 0; k < nsteps; k++) {
    if (s->solsteps < 1 || s->solsteps > ISPH_MAX_SOLSTEPS) return ISPH_ENONFINITE;
    const double told = s->t;
    s->t = s->t + s->dt;
    for (int q = 0; q < s->solsteps; q++) {
      or_scalart(s);
      or_scalarupdate(s, s->dtsol);
    }
    int brc = buffer_clock(s, told);
    if (brc) return brc;
    or_projection(s);
    if (s->wave) paddle_to(s, s->t);
    or_cellshift(s);
    int rc = or_neighbour(s);
    if (rc) return rc;
    or_effpor(s);
    or_eddyvis(s);
    or_int_vel(s);
    if (s->wcsph) {
      /* weakly compressible form: continuity and the Tait law in place of
       * the pressure Poisson solve */
      or_calmasscons(s);
      or_updateden(s);
      or_dentopress(s);
      if (st) memset(st, 0, sizeof *st);
    } else {
      rc = or_ppe_assemble(s);
      if (rc) return rc;
      rc = or_ppe_solve(s, st);
      if (rc == ISPH_ENONFINITE) return rc;
      if (rc) rcs = rc;
    }
    or_comp_vel(s);
    or_comp_pos(s);
    or_cellshiftalt(s);
    rc = or_neighbour(s);
    if (rc) return rc;
    or_freesurf(s);
    or_effpor(s);
    or_opt2_shift(s);
    if (!s->wcsph) or_densityupdate(s);
    or_timestep(s);
    or_eddyvis(s);
  }
  return rcs;
}

/* ------------------------------------------------------------------ */
/* accessors for the ctypes wrapper                                     */
/* ------------------------------------------------------------------ */
void *or_array(ostate *s, const char *name) {
  struct { const char *n; void *p; } t[] = {
      {"x", s->x}, {"y", s->y}, {"u", s->u}, {"v", s->v}, {"us", s->us}, {"vs", s->vs},
      {"xs", s->xs}, {"ys", s->ys}, {"p", s->p}, {"m", s->m}, {"rho", s->rho},
      {"rho0", s->rho0}, {"vol0", s->vol0}, {"c", s->c}, {"tid", s->tid}, {"cell", s->cell},
      {"flags", s->flags}, {"por", s->por}, {"resist", s->resist}, {"gradvx", s->gradvx},
      {"coff", s->coff}, {"nut", s->nut}, {"tke", s->tke}, {"varts", s->varts},
      {"cdiff", s->cdiff}, {"vel", s->vel}, {"ncnt", s->ncnt}, {"nbr", s->nbr}, {"nr", s->nr},
      {"klt1", s->klt1}, {"klt2", s->klt2}, {"cstart", s->cstart}, {"alen", s->alen},
      {"acol", s->acol}, {"aval", s->aval}, {"b", s->b}, {"sol", s->sol}, {"sx", s->sx},
      {"sy", s->sy}, {"sm", s->sm}, {"srho", s->srho}, {"sstart", s->sstart}, {"tden", s->tden},
      {"cmaxvel", s->cmaxvel}, {"cmaxeddy", s->cmaxeddy}};
  for (size_t k = 0; k < sizeof(t) / sizeof(t[0]); k++)
    if (strcmp(t[k].n, name) == 0) return t[k].p;
  return NULL;
}

void or_get_scalars(const ostate *s, isph_scalars *o) {
  o->t = s->t; o->dt = s->dt; o->dtsol = s->dtsol; o->umax = s->umax; o->numax = s->numax;
  o->solsteps = s->solsteps; o->n = (int32_t)s->n;
}

void or_set_scalars(ostate *s, const isph_scalars *o) {
  s->t = o->t; s->dt = o->dt; s->dtsol = o->dtsol; s->umax = o->umax; s->numax = o->numax;
  s->solsteps = o->solsteps;
}

/* Kernel-function probes for the known-answer tests. */
double or_wab(double r, double h) { return o_wab(r, h); }
double or_wo(double h) { return o_wo(h); }
double or_wabd(double d, double r, double h) { return o_wabd(d, r, h); }
void or_inv2(double *a) { o_inv2(a); }
double or_vis(double oden) { return o_vis(oden); }
double or_reduce(int64_t n, const double *prod) { return reduce_prod(NULL, n, prod); }
int or_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Standalone dense-free solve of an explicit system for the KAT: rows given
 * in canonical ELL form. Uses the same BiCGStab as the PPE path. */
int or_solve_system(ostate *s, int64_t n, const int32_t *len, const int32_t *col,
                    const double *val, const double *rhs, double *x, isph_solve_stats *st) {
  if (n > s->cap) return ISPH_ECAP;
  s->n = n;
  for (int64_t i = 0; i < n; i++) {
    s->alen[i] = len[i];
    for (int k = 0; k < len[i]; k++) {
      s->acol[i * RCAP + k] = col[i * RCAP + k];
      s->aval[i * RCAP + k] = val[i * RCAP + k];
    }
    s->b[i] = rhs[i];
    s->sol[i] = x[i];
  }
  s->assembled = 1;
  int rc = or_ppe_solve(s, st);
  for (int64_t i = 0; i < n; i++) x[i] = s->sol[i];
  return rc;
}
