// isph_lib.hip -- C ABI of the MI355X ISPH per-step pipeline (declared in
// include/isph_abi.h).  One context = one scenario, device resident, one HIP
// stream.  Every entry point maps to one reference subroutine of the
// production loop (main.f90:199-261) or to a fused sequence of them.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/isph_abi.h"
#include "isph_device.h"
#include "isph_kernels.h"
#include "isph_phases.h"
#include "isph_solver.h"
#include "isph_mg.h"
#include "isph_batch.h"
#include "isph_buffer.h"
#include "isph_wcsph.h"

static_assert(NBR_CAP == ISPH_NBR_CAP, "neighbour capacity mismatch");

// HIP-event timing of one kernel class
struct KTimer {
  const char* name;
  int per_iter;  // launches per solver iteration (the working launches of a chunk)
  double total_ms = 0;
  int64_t launches = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool, pending;
};

// Kernels whose block size the autotuner chooses, in the order of
// isph_autotune's output (their results do not depend on it: one thread per
// particle, and only integer or max reductions across a block).
enum {
  TK_NEIGHBOUR, TK_EFFPOR, TK_EDDYVIS, TK_VARTS, TK_INTVEL, TK_ASSEMBLE, TK_COMPVEL,
  TK_SHIFT, TK_SCALART, TK_N
};
static_assert(TK_N == ISPH_TUNE_N, "tuned kernel count mismatch");
// launch shapes the tuner chooses from: particles in blocks of 64, 128 or
// 256 threads; for the neighbour search also cell tiles of 4 rows on 128
// threads or 8 rows on 256 (LDS-staged)
enum { TS_P64, TS_P128, TS_P256, TS_TILE4, TS_TILE8, TS_N };
static_assert(TS_N == ISPH_TUNE_C, "launch shape count mismatch");
static const int ts_block[TS_N] = {64, 128, 256, 128, 256};
static int ts_count(int tk) { return tk == TK_NEIGHBOUR ? TS_N : TS_TILE4; }

// an inflow/outflow buffer zone: device copies of the template and its cell
// tables, and the insertion clock (main.f90:217-235)
struct HostZone {
  bool set = false;
  int nt = 0;
  double *tx = nullptr, *ty = nullptr, *tu = nullptr, *tv = nullptr, *tp = nullptr;
  int32_t *tcell = nullptr, *trank = nullptr, *tcnt = nullptr;
  uint8_t* entry = nullptr;
  double mass = 0, rho = 0, vol0 = 0, c = 0, rv = 0, interval = 0, time_shift = 0;
  int iter = 1;
};

struct isph_ctx {
  isph_params P;
  HostZone zones[2];
  bool wcsph = false;      // fused step in the weakly compressible form
  // piston wavemaker: stroke, period, ramp time and start time of the paddle motion
  bool wave = false;
  double wv_stroke = 0.0, wv_period = 1.0, wv_ramp = 0.0, wv_t0 = 0.0;
  bool prof = false;
  // per_iter of mg_down and mg_up (2 x the multi-workgroup cell levels) is
  // set by each solve
  KTimer timers[11] = {{"spmv_p", 1},    {"spmv_s", 1},  {"spmv_p32", 1},     {"spmv_s32", 1},
                       {"mg_coarse", 2}, {"mg_pre0", 2}, {"mg_post0", 2},     {"mg_down", 0},
                       {"mg_up", 0},     {"pfinalize", 2}, {"mg_restrict0", 2}};
  Dev D;
  hipStream_t stream = nullptr;
  int64_t cap = 0, scap = 0;
  int n_hi = 0;          // host-side upper bound of the live particle count
  int ncell = 0;
  int device = 0;
  bool assembled = false, have_nbr = false;
  isph_solve_stats last{};
  std::string err;
  std::vector<void*> allocs;
  double* d_bn2 = nullptr;
  MgDev G{};               // multigrid hierarchy (precision form)
  struct BatchHost* self_batch = nullptr;  // this context as a batch of one (precision solve)
  int co_lds = 0;                          // LDS bytes of its coarse multigrid workgroup
  // previous precision solve: iterations to the first stop, and after it
  // (chunk sizing only; the results do not depend on the chunks)
  int pred_first = 0, pred_after = 0;
  bool in_batch = false;   // a member of a batch: its form (precision, WCSPH) is fixed
  int tshape[TK_N] = {TS_P256, TS_P256, TS_P256, TS_P256, TS_P256,
                     TS_P256, TS_P256, TS_P256, TS_P256};  // tuned launch shapes
};

#define HIPCHK(ctx, expr)                                                     \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      set_err(ctx, "%s failed: %s", #expr, hipGetErrorString(e_));            \
      return ISPH_EHIP;                                                       \
    }                                                                         \
  } while (0)

static void set_err(isph_ctx* c, const char* fmt, ...) {
  if (!c) return;
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  c->err = buf;
}

// message of the calling thread's last failed isph_create or
// isph_ppe_solve_noalloc (no context exists to hold it); returned by
// isph_last_error(NULL)
static thread_local std::string g_noctx_err;

static int create_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_noctx_err = buf;
  return code;
}

static int grid_of(int n) { return std::max(1, (n + 255) / 256); }
// grid of a row kernel mapped by row_block(): a multiple of the 8 XCDs
static int row_grid(int n) { return 8 * ((grid_of(n) + 7) / 8); }
// the host's upper bound of the live particle count and the row blocks over it
static void set_nhi(isph_ctx* c, int n) {
  c->n_hi = n;
  c->D.nblk = grid_of(n);
}

template <typename T>
static int dalloc(isph_ctx* c, T** p, int64_t count) {
  void* q = nullptr;
  size_t bytes = sizeof(T) * (size_t)std::max<int64_t>(count, 1);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    set_err(c, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return ISPH_EHIP;
  }
  c->allocs.push_back(q);
  *p = (T*)q;
  return 0;
}

// wab on the host, identical expression to the device (kernel.f90:34-92)
static double host_wo(double h, double k1, double k2) {
  return ((0.1 * 4.0) * k1) + (((1.0 - 0.1) * 6.0) * k2);
}

static void fill_devparams(isph_ctx* c) {
  const isph_params& P = c->P;
  DevParams& d = c->D.P;
  d.cellx = P.cellx;
  d.celly = P.celly;
  d.h = P.h; d.dl = P.dl; d.lam = P.lam; d.rho = P.rho; d.rel_den = P.rel_den;
  d.rhomin = P.rhomin; d.g = P.g; d.mu = P.mu; d.por = P.por; d.d50 = P.d50;
  d.lamk = P.lamk; d.csh = P.csh; d.maxshift = P.maxshift; d.fs_thresh = P.fs_thresh;
  d.sig1 = P.sig1; d.tschmidt = P.tschmidt; d.xlcutoff = P.xlcutoff; d.ytcutoff = P.ytcutoff;
  d.buf_xmin = P.buf_xmin; d.buf_xmax = P.buf_xmax; d.tol = P.tol; d.inner_tol = P.inner_tol;
  d.fixed_dt = P.fixed_dt;
  d.max_iter = P.max_iter;
  d.max_outer = P.max_outer;
  d.cs = 4.0 * P.h;
  d.k1 = 10.0 / ((double)(28.0f * 3.14f) * (P.h * P.h));
  d.k2 = 1.0 / ((double)(3.0f * 3.14f) * (P.h * P.h));
  d.two_h = 2 * P.h;
  d.lim2 = (d.two_h * d.two_h) * (1.0 + 1e-9);
  d.wo = host_wo(P.h, d.k1, d.k2);
}

extern "C" int64_t isph_footprint_bytes(int64_t n_cap, int64_t n_solid_cap, int32_t cellx,
                                        int32_t celly, int32_t precision) {
  const int64_t ncell = (int64_t)cellx * celly;
  int64_t per = 0;
  per += 2 * (14 * 8 + 3 + 4);                  // gathered state, ping-pong
  per += 14 * 8;                                // derived properties
  per += 4 + 4 * NBR_CAP;                       // neighbour list
  per += 4 * 3;                                 // bin, rank, scan scratch
  per += 4 + ROW_CAP * (4 + 8) + 2 * 8;         // PPE rows, rhs, solution
  per += 6 * 8;                                 // FP64 solver vectors
  // precision form: FP16 V-cycle matrix and packed column offsets, FP32
  // vectors, level-0 multigrid vectors, setup rows
  if (precision) per += 2 * 4 * ((ROW_CAP + 1) / 2) + 5 * 4 + 2 * 4 + 9 * 8;
  int64_t b = per * n_cap + 4 * 8 * n_solid_cap + 4 * (ncell * (2 + 2 * NCLS) + 16);
  b += 5 * 8 * (n_cap / 256 + 2);               // reduction partials
  b += 8 * (2 * ncell + 2 * (ncell / 256 + 2)); // per-cell maxima and their block maxima
  // cell levels: stencil, x, b, r and 1/diagonal per unknown, the levels
  // holding 4/3 of the cell count
  if (precision) b += 4 * 13 * ((4 * ncell) / 3 + MG_MAX_LEVELS) + 8 * MG_DENSE_MAX * MG_DENSE_MAX;
  return b + (1 << 20);
}

extern "C" int isph_create(const isph_params* params, int64_t n_cap, int64_t n_solid_cap,
                           isph_ctx** out) {
  // slot-major ELL indices k*cap + i are int32: cap*ROW_CAP must stay below 2^31
  if (out) *out = nullptr;
  g_noctx_err.clear();
  if (!params || !out) return create_fail(ISPH_EARG, "isph_create: null argument");
  if (n_cap <= 0 || n_cap * ROW_CAP >= (1ll << 31) || n_solid_cap < 0)
    return create_fail(ISPH_EARG, "isph_create: capacity %lld outside (0, 2^31/%d)",
                       (long long)n_cap, ROW_CAP);
  if (params->cellx < 3 || params->celly < 3 || !(params->h > 0))
    return create_fail(ISPH_EARG, "isph_create: cell grid %d x %d or h %g invalid",
                       params->cellx, params->celly, params->h);
  if ((int64_t)params->cellx * params->celly * NCLS + 1 > (1ll << 30))
    return create_fail(ISPH_EARG, "isph_create: cell grid too large");
  isph_ctx* c = new isph_ctx();
  c->P = *params;
  c->cap = n_cap;
  c->scap = std::max<int64_t>(n_solid_cap, 1);
  c->ncell = params->cellx * params->celly;
  hipError_t he = hipGetDevice(&c->device);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (he != hipSuccess) {
    delete c;
    return create_fail(ISPH_EHIP, "isph_create: no usable HIP device: %s", hipGetErrorString(he));
  }
  Dev& D = c->D;
  D.cap = (int)n_cap;
  fill_devparams(c);
  int rc = 0;
  const int64_t N = n_cap;
  double** dd[] = {&D.x, &D.y, &D.u, &D.v, &D.us, &D.vs, &D.xs, &D.ys, &D.p, &D.m, &D.rho,
                   &D.rho0, &D.vol0, &D.c, &D.x2, &D.y2, &D.u2, &D.v2, &D.us2, &D.vs2,
                   &D.xs2, &D.ys2, &D.p2, &D.m2, &D.rho2, &D.rho02, &D.vol02, &D.c2,
                   &D.por, &D.resist, &D.gradvx, &D.cf1, &D.cf2, &D.cf3, &D.cf4, &D.nut,
                   &D.tke, &D.varts, &D.cdiff, &D.vel, &D.xnb, &D.ynb, &D.b, &D.sol,
                   &D.r, &D.rh, &D.pp, &D.t1, &D.sv, &D.t2, &D.tden};
  for (auto p : dd) rc |= dalloc(c, p, N);
  uint8_t** bb[] = {&D.tid, &D.flags, &D.vic, &D.tid2, &D.flags2, &D.vic2};
  for (auto p : bb) rc |= dalloc(c, p, N);
  rc |= dalloc(c, &D.cell, N);
  rc |= dalloc(c, &D.cell2, N);
  rc |= dalloc(c, &D.ncnt, N);
  rc |= dalloc(c, &D.nbr, N * NBR_CAP);
  rc |= dalloc(c, &D.pbin, N);
  rc |= dalloc(c, &D.prank, N);
  rc |= dalloc(c, &D.alen, N);
  rc |= dalloc(c, &D.acol, N * ROW_CAP);
  rc |= dalloc(c, &D.aval, N * ROW_CAP);
  const int64_t nbins = (int64_t)c->ncell * NCLS + 1;
  rc |= dalloc(c, &D.cstart, c->ncell + 1);
  rc |= dalloc(c, &D.bincnt, nbins);
  rc |= dalloc(c, &D.binstart, nbins);
  rc |= dalloc(c, &D.scan_tmp, nbins / SCAN_TILE + 2);
  rc |= dalloc(c, &D.sx, c->scap);
  rc |= dalloc(c, &D.sy, c->scap);
  rc |= dalloc(c, &D.sm, c->scap);
  rc |= dalloc(c, &D.srho, c->scap);
  rc |= dalloc(c, &D.sstart, c->ncell + 1);
  rc |= dalloc(c, &D.part, 5 * (N / 256 + 2));
  D.nmaxb = (int)(c->ncell / 256 + 2);
  rc |= dalloc(c, &D.maxb, 2 * (int64_t)D.nmaxb);
  rc |= dalloc(c, &D.cmaxv, c->ncell);
  rc |= dalloc(c, &D.cmaxn, c->ncell);
  // the per-cell maxima start at zero (an empty cell keeps them until it fills)
  if (rc == 0 && (hipMemset(D.maxb, 0, sizeof(unsigned long long) * 2 * (size_t)D.nmaxb) !=
                      hipSuccess ||
                  hipMemset(D.cmaxv, 0, sizeof(unsigned long long) * (size_t)c->ncell) !=
                      hipSuccess ||
                  hipMemset(D.cmaxn, 0, sizeof(unsigned long long) * (size_t)c->ncell) !=
                      hipSuccess))
    rc |= ISPH_EHIP;
  rc |= dalloc(c, &D.sc, 1);
  rc |= dalloc(c, &c->d_bn2, 1);
  if (params->precision) {
    rc |= dalloc(c, &D.av16, N * ((ROW_CAP + 1) / 2));
    rc |= dalloc(c, &D.aoff, N * ((ROW_CAP + 1) / 2));
    float** ff[] = {&D.r32, &D.rh32, &D.pp32, &D.sv32, &D.d32};
    for (auto p : ff) rc |= dalloc(c, p, N);
    // cell hierarchy: level 1 is the cell grid, then 2x2 aggregation down to
    // at most MG_DENSE_MAX unknowns
    MgDev& G = c->G;
    int ny = params->celly, nx = params->cellx, off = 0, l = 1;
    for (;;) {
      if (l >= MG_MAX_LEVELS) { rc |= ISPH_EARG; break; }
      G.lev[l].ny = ny; G.lev[l].nx = nx; G.lev[l].n = ny * nx; G.lev[l].off = off;
      off += ny * nx;
      if (ny * nx <= MG_DENSE_MAX) break;
      ny = (ny + 1) / 2; nx = (nx + 1) / 2; l++;
    }
    G.nlev = l;
    rc |= dalloc(c, &G.sten, 9 * (int64_t)off);
    rc |= dalloc(c, &G.xb, off);
    rc |= dalloc(c, &G.bb, off);
    rc |= dalloc(c, &G.rb, off);
    rc |= dalloc(c, &G.rdg, off);
    rc |= dalloc(c, &G.dense_inv, MG_DENSE_MAX * MG_DENSE_MAX);
    rc |= dalloc(c, &G.rowst, 9 * N);
    rc |= dalloc(c, &G.d0, N);
    rc |= dalloc(c, &G.x0, N);
    G.w = MG_OMEGA;
    G.cscale = MG_CSCALE;
  }
  if (rc) {
    const std::string e = c->err;
    isph_destroy(c);
    return create_fail(ISPH_EHIP, "isph_create: %s", e.c_str());
  }
  c->G.done = &D.sc->done;
  if (params->precision) {
    // The coarse workgroup runs the longest tail of levels l0..nlev whose
    // first level has at most MG_CO_PER unknowns per thread and whose data
    // fits the LDS budget; levels 1..l0-1 run multi-workgroup.
    MgDev& G = c->G;
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
        hipSuccess)
      lds_max = 64 * 1024;
    int64_t budget = std::min<int64_t>(lds_max, MG_CO_LDS_MAX) - 1024;
    // dynamic LDS above 64 KB is requested explicitly; if the runtime
    // refuses, the coarse tail is sized for 64 KB
    if (budget > 63 * 1024 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&kb_mg_coarse),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)budget) != hipSuccess) {
      (void)hipGetLastError();
      budget = 63 * 1024;
    }
    int l0 = 1;
    while (l0 < G.nlev && (G.lev[l0].n > MG_CO_PER * MG_CO_THREADS ||
                           mg_coarse_lds_bytes(G, l0) > budget))
      l0++;
    G.nmb = l0 - 1;
    c->co_lds = (int)mg_coarse_lds_bytes(G, l0);
  }
  DevScalars s{};
  s.dt = params->dt;
  s.dtsol = params->dt;
  s.solsteps = 1;
  he = hipMemcpy(D.sc, &s, sizeof s, hipMemcpyHostToDevice);
  if (he != hipSuccess) {
    isph_destroy(c);
    return create_fail(ISPH_EHIP, "isph_create: hipMemcpy failed: %s", hipGetErrorString(he));
  }
  (void)hipMemset(D.sstart, 0, sizeof(int32_t) * (c->ncell + 1));
  (void)hipMemset(D.vic, 0, N);
  *out = c;
  return 0;
}

static void bh_free(BatchHost* b);

extern "C" int isph_destroy(isph_ctx* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  bh_free(c->self_batch);
  for (HostZone& z : c->zones) {
    void* zp[] = {z.tx, z.ty, z.tu, z.tv, z.tp, z.tcell, z.trank, z.tcnt, z.entry};
    for (void* q : zp) (void)hipFree(q);
  }
  for (void* p : c->allocs) (void)hipFree(p);
  for (KTimer& t : c->timers) {
    for (auto& pr : t.pool) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    for (auto& pr : t.pending) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

extern "C" const char* isph_last_error(const isph_ctx* c) {
  if (c) return c->err.c_str();
  return g_noctx_err.empty() ? "null context" : g_noctx_err.c_str();
}

extern "C" int isph_set_params(isph_ctx* c, const isph_params* p) {
  if (!c || !p) return ISPH_EARG;
  if (p->cellx != c->P.cellx || p->celly != c->P.celly) return ISPH_EARG;
  if (c->in_batch && p->precision != c->P.precision) {
    set_err(c, "the solve form of a context cannot change while it belongs to a batch");
    return ISPH_ESTATE;
  }
  if (p->precision && !c->D.av16) {
    set_err(c, "context was created without the precision-form buffers");
    return ISPH_EARG;
  }
  c->P = *p;
  fill_devparams(c);
  return 0;
}

// containing interior cell with half-open bounds (internal.f90:80-83)
static int locate_cell(const isph_params& P, double x, double y) {
  const double cs = 4.0 * P.h;
  if (!(x == x) || !(y == y)) return -1;
  int cj = (int)std::floor(x / cs) + 1;
  int ci = P.celly - 2 - (int)std::floor(y / cs);
  auto xl = [&](int j) { return (double)(j - 1) * cs; };
  auto xr = [&](int j) { return (double)(j)*cs; };
  auto yb = [&](int i) { return (double)(P.celly - 2 - i) * cs; };
  auto yt = [&](int i) { return (double)(P.celly - 1 - i) * cs; };
  for (int it = 0; it < 3; it++) {
    if (cj < 0 || cj >= P.cellx) break;
    if (x < xl(cj)) cj--;
    else if (x >= xr(cj)) cj++;
  }
  for (int it = 0; it < 3; it++) {
    if (ci < 0 || ci >= P.celly) break;
    if (y < yb(ci)) ci++;
    else if (y >= yt(ci)) ci--;
  }
  if (cj < 1 || cj > P.cellx - 2 || ci < 1 || ci > P.celly - 2) return -1;
  if (!(x >= xl(cj) && x < xr(cj) && y >= yb(ci) && y < yt(ci))) return -1;
  return cj * P.celly + ci;
}

static int sync(isph_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

static int read_scalars(isph_ctx* c, DevScalars* s) {
  HIPCHK(c, hipMemcpyAsync(s, c->D.sc, sizeof *s, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int isph_upload(isph_ctx* c, int64_t n, const double* x, const double* y,
                           const double* u, const double* v, const double* p, const double* m,
                           const double* rho, const double* rho0, const double* vol0,
                           const double* cc, const int32_t* tid, const uint8_t* flags) {
  if (!c) return ISPH_EARG;
  if (n < 0 || n > c->cap) {
    set_err(c, "upload of %lld particles exceeds capacity %lld", (long long)n, (long long)c->cap);
    return ISPH_ECAP;
  }
  if (n > 0 && (!x || !y || !u || !v || !p || !m || !rho || !rho0 || !vol0 || !cc || !tid || !flags))
    return ISPH_EARG;
  // kernels still in flight on the context's stream read the arrays written here
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // initial binning: walls before fluid in each cell, input order kept
  std::vector<int64_t> key(n);
  for (int64_t i = 0; i < n; i++) {
    int k = locate_cell(c->P, x[i], y[i]);
    if (k < 0) {
      set_err(c, "particle %lld at (%g, %g) is outside the interior cells", (long long)i, x[i], y[i]);
      return ISPH_EARG;
    }
    if (tid[i] < 1 || tid[i] > 3) {
      set_err(c, "particle %lld has tid %d", (long long)i, tid[i]);
      return ISPH_EARG;
    }
    if (tid[i] == ISPH_TID_FLUID && (flags[i] & ISPH_F_PADDLE)) {
      set_err(c, "particle %lld is fluid and flagged as a paddle wall", (long long)i);
      return ISPH_EARG;
    }
    key[i] = 2 * (int64_t)k + (tid[i] == ISPH_TID_FLUID ? 1 : 0);
  }
  std::vector<int64_t> ord(n);
  for (int64_t i = 0; i < n; i++) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  std::vector<double> buf(n);
  auto put = [&](double* dst, const double* src) -> int {
    for (int64_t k = 0; k < n; k++) buf[k] = src[ord[k]];
    HIPCHK(c, hipMemcpy(dst, buf.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    return 0;
  };
  Dev& D = c->D;
  int rc = 0;
  rc |= put(D.x, x); rc |= put(D.y, y); rc |= put(D.u, u); rc |= put(D.v, v); rc |= put(D.p, p);
  rc |= put(D.m, m); rc |= put(D.rho, rho); rc |= put(D.rho0, rho0); rc |= put(D.vol0, vol0);
  rc |= put(D.c, cc);
  rc |= put(D.sol, p);  // fsol starts from the initial pressures (memory.f90:113-144)
  if (rc) return rc;
  std::vector<double> zeros(n, 0.0);
  HIPCHK(c, hipMemcpy(D.us, zeros.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(D.vs, zeros.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(D.xs, zeros.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(D.ys, zeros.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  std::vector<uint8_t> tb(n), fb(n), vb(n);
  std::vector<int32_t> cb(n);
  for (int64_t k = 0; k < n; k++) {
    tb[k] = (uint8_t)tid[ord[k]];
    fb[k] = flags[ord[k]] & (ISPH_F_BUFFER | ISPH_F_FREE | ISPH_F_PADDLE);
    vb[k] = (flags[ord[k]] & ISPH_F_VICINITY) ? 1 : 0;
    cb[k] = (int32_t)(key[ord[k]] / 2);
  }
  HIPCHK(c, hipMemcpy(D.tid, tb.data(), n, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(D.flags, fb.data(), n, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(D.vic, vb.data(), n, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(D.cell, cb.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  // porosity defaults to 1 (particle.f90:62-64)
  std::vector<double> ones(n, 1.0);
  HIPCHK(c, hipMemcpy(D.por, ones.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  std::vector<int32_t> cs(c->ncell + 1, 0);
  for (int64_t k = 0; k < n; k++) cs[cb[k] + 1]++;
  for (int k = 0; k < c->ncell; k++) cs[k + 1] += cs[k];
  HIPCHK(c, hipMemcpy(D.cstart, cs.data(), sizeof(int32_t) * (c->ncell + 1), hipMemcpyHostToDevice));
  DevScalars s;
  HIPCHK(c, hipMemcpy(&s, D.sc, sizeof s, hipMemcpyDeviceToHost));
  s.n = (int)n;
  s.err = 0;
  s.umax_bits = 0;
  s.numax_bits = 0;
  HIPCHK(c, hipMemcpy(D.sc, &s, sizeof s, hipMemcpyHostToDevice));
  set_nhi(c, (int)n);
  c->assembled = false;
  c->have_nbr = false;
  return 0;
}

extern "C" int isph_upload_solids(isph_ctx* c, int64_t ns, const double* x, const double* y,
                                  const double* m, const double* rho) {
  if (!c) return ISPH_EARG;
  if (ns < 0 || ns > c->scap) {
    set_err(c, "upload of %lld solids exceeds capacity %lld", (long long)ns, (long long)c->scap);
    return ISPH_ECAP;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int32_t> key(ns);
  std::vector<int32_t> cs(c->ncell + 1, 0);
  for (int64_t i = 0; i < ns; i++) {
    int k = locate_cell(c->P, x[i], y[i]);
    if (k < 0) {
      set_err(c, "solid %lld at (%g, %g) is outside the interior cells", (long long)i, x[i], y[i]);
      return ISPH_EARG;
    }
    key[i] = k;
    cs[k + 1]++;
  }
  for (int k = 0; k < c->ncell; k++) cs[k + 1] += cs[k];
  std::vector<int32_t> fill(cs.begin(), cs.end());
  std::vector<double> bx(ns), by(ns), bm(ns), br(ns);
  for (int64_t i = 0; i < ns; i++) {
    int64_t d = fill[key[i]]++;
    bx[d] = x[i]; by[d] = y[i]; bm[d] = m[i]; br[d] = rho[i];
  }
  Dev& D = c->D;
  if (ns) {
    HIPCHK(c, hipMemcpy(D.sx, bx.data(), 8 * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(D.sy, by.data(), 8 * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(D.sm, bm.data(), 8 * ns, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(D.srho, br.data(), 8 * ns, hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipMemcpy(D.sstart, cs.data(), 4 * (c->ncell + 1), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int isph_download(isph_ctx* c, int64_t* n_out, double* x, double* y, double* u,
                             double* v, double* p, double* m, double* rho, double* rho0,
                             double* vol0, double* cc, int32_t* tid, uint8_t* flags) {
  if (!c) return ISPH_EARG;
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  const int64_t n = s.n;
  set_nhi(c, s.n);
  if (n_out) *n_out = n;
  Dev& D = c->D;
  struct { double* dst; double* src; } f[] = {{x, D.x}, {y, D.y}, {u, D.u}, {v, D.v}, {p, D.p},
                                              {m, D.m}, {rho, D.rho}, {rho0, D.rho0},
                                              {vol0, D.vol0}, {cc, D.c}};
  for (auto& e : f)
    if (e.dst) HIPCHK(c, hipMemcpy(e.dst, e.src, 8 * n, hipMemcpyDeviceToHost));
  if (tid) {
    std::vector<uint8_t> t(n);
    HIPCHK(c, hipMemcpy(t.data(), D.tid, n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) tid[i] = t[i];
  }
  if (flags) {
    std::vector<uint8_t> fl(n), vi(n);
    HIPCHK(c, hipMemcpy(fl.data(), D.flags, n, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(vi.data(), D.vic, n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) flags[i] = fl[i] | (vi[i] ? ISPH_F_VICINITY : 0);
  }
  return 0;
}

extern "C" int isph_get_scalars(isph_ctx* c, isph_scalars* o) {
  if (!c || !o) return ISPH_EARG;
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  o->t = s.t; o->dt = s.dt; o->dtsol = s.dtsol; o->umax = s.umax; o->numax = s.numax;
  o->solsteps = s.solsteps; o->n = s.n;
  return 0;
}

extern "C" int isph_set_scalars(isph_ctx* c, const isph_scalars* in) {
  if (!c || !in) return ISPH_EARG;
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  s.t = in->t; s.dt = in->dt; s.dtsol = in->dtsol; s.umax = in->umax; s.numax = in->numax;
  s.solsteps = in->solsteps;
  HIPCHK(c, hipMemcpy(c->D.sc, &s, sizeof s, hipMemcpyHostToDevice));
  return 0;
}

// ---------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------
#define LAUNCH(kern, nthreads, ...)                                                  \
  do {                                                                               \
    hipLaunchKernelGGL(kern, dim3(grid_of(nthreads)), dim3(256), 0, c->stream, __VA_ARGS__); \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess) {                                                          \
      set_err(c, "launch of %s failed: %s", #kern, hipGetErrorString(e_));          \
      return ISPH_EHIP;                                                              \
    }                                                                                \
  } while (0)

// launch of a row kernel mapped by row_block() (nthreads = c->n_hi)
#define RLAUNCH(kern, nthreads, ...)                                                 \
  do {                                                                               \
    hipLaunchKernelGGL(kern, dim3(row_grid(nthreads)), dim3(256), 0, c->stream, __VA_ARGS__); \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess) {                                                          \
      set_err(c, "launch of %s failed: %s", #kern, hipGetErrorString(e_));          \
      return ISPH_EHIP;                                                              \
    }                                                                                \
  } while (0)

// launch of a tuned particle kernel with the context's block size for it
#define TLAUNCH(tk, kern, nthreads, ...)                                             \
  do {                                                                               \
    const int b_ = ts_block[c->tshape[tk]];                                          \
    hipLaunchKernelGGL(kern, dim3(std::max(1, ((nthreads) + b_ - 1) / b_)), dim3(b_), 0, \
                       c->stream, __VA_ARGS__);                                      \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess) {                                                          \
      set_err(c, "launch of %s failed: %s", #kern, hipGetErrorString(e_));          \
      return ISPH_EHIP;                                                              \
    }                                                                                \
  } while (0)

static std::pair<hipEvent_t, hipEvent_t> timer_pair(KTimer& t) {
  if (!t.pool.empty()) {
    auto p = t.pool.back();
    t.pool.pop_back();
    return p;
  }
  std::pair<hipEvent_t, hipEvent_t> p;
  (void)hipEventCreate(&p.first);
  (void)hipEventCreate(&p.second);
  return p;
}

// Collect the elapsed times of completed launches (called after a sync).
// Only the launches of the first `keep` pending iterations of each timer are
// counted (per_iter launches each): the solver launches whole chunks of
// iterations, and the launches after the device has set its done flag
// return at once; they are dropped so that the average is over launches
// that did the full work.  keep < 0 counts all.
static void timers_collect(isph_ctx* c, int64_t keep) {
  for (KTimer& t : c->timers) {
    int64_t k = 0;
    for (auto& pr : t.pending) {
      float ms = 0.f;
      if ((keep < 0 || k < keep * t.per_iter) &&
          hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
        t.total_ms += ms;
        t.launches += 1;
      }
      k++;
      t.pool.push_back(pr);
    }
    t.pending.clear();
  }
}

// A timed launch of a row kernel: the two events are attached to the kernel's own dispatch
// (hipExtLaunchKernelGGL), so their difference is the kernel's execution
// time without the gaps around it.
#define TIMED_LAUNCH(tidx, kern, nthreads, ...)                                      \
  do {                                                                               \
    if (c->prof) {                                                                   \
      std::pair<hipEvent_t, hipEvent_t> ev_ = timer_pair(c->timers[tidx]);           \
      hipExtLaunchKernelGGL(kern, dim3(row_grid(nthreads)), dim3(256), 0, c->stream, \
                            ev_.first, ev_.second, 0, __VA_ARGS__);                  \
      hipError_t e_ = hipGetLastError();                                             \
      if (e_ != hipSuccess) {                                                        \
        set_err(c, "launch of %s failed: %s", #kern, hipGetErrorString(e_));        \
        return ISPH_EHIP;                                                            \
      }                                                                              \
      c->timers[tidx].pending.push_back(ev_);                                        \
    } else {                                                                         \
      RLAUNCH(kern, nthreads, __VA_ARGS__);                                          \
    }                                                                                \
  } while (0)

extern "C" int isph_prof_enable(isph_ctx* c, int32_t on) {
  if (!c) return ISPH_EARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  timers_collect(c, -1);
  c->prof = on != 0;
  for (KTimer& t : c->timers) {
    t.total_ms = 0;
    t.launches = 0;
  }
  return 0;
}

extern "C" int isph_prof_read(isph_ctx* c, const char* kernel, double* total_ms,
                              int64_t* launches) {
  if (!c || !kernel) return ISPH_EARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  timers_collect(c, -1);
  for (KTimer& t : c->timers) {
    if (std::strcmp(t.name, kernel) == 0) {
      if (total_ms) *total_ms = t.total_ms;
      if (launches) *launches = t.launches;
      return 0;
    }
  }
  return ISPH_EARG;
}

// one-block reduction of the partials written by the preceding row kernel
#define FINALIZE_OUT(op, nrows, outp)                                                \
  do {                                                                               \
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, c->stream, c->D, (op),     \
                       grid_of(nrows), (outp));                                      \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess) {                                                          \
      set_err(c, "launch of k_finalize failed: %s", hipGetErrorString(e_));         \
      return ISPH_EHIP;                                                              \
    }                                                                                \
  } while (0)
#define FINALIZE(op, nrows) FINALIZE_OUT(op, nrows, (double*)nullptr)

static int check_err_flags(isph_ctx* c) {
  int e = 0;
  HIPCHK(c, hipMemcpyAsync(&e, &c->D.sc->err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (e) {
    int z = 0;
    HIPCHK(c, hipMemcpy(&c->D.sc->err, &z, sizeof z, hipMemcpyHostToDevice));
    if (e & ERR_NBR_CAP) {
      set_err(c, "a particle has more than %d neighbours", NBR_CAP);
      return ISPH_ECAP;
    }
    if (e & ERR_DIAG) {
      set_err(c, "zero or non-finite PPE diagonal (isolated particle)");
      return ISPH_ENONFINITE;
    }
  }
  return 0;
}

extern "C" int isph_projection(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  LAUNCH(k_projection, c->n_hi, c->D);
  return 0;
}

static void swap_state(Dev& D) { dev_swap_state(D); }

static int scan_bins(isph_ctx* c, int len) {
  const int ntiles = (len + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(256), 0, c->stream, c->D.bincnt,
                     c->D.binstart, c->D.scan_tmp, len);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, c->stream, c->D.scan_tmp, ntiles,
                     (int32_t*)nullptr);
  hipLaunchKernelGGL(k_scan_add, dim3(grid_of(len)), dim3(256), 0, c->stream, c->D.binstart,
                     c->D.scan_tmp, len);
  HIPCHK(c, hipGetLastError());
  return 0;
}

static int rebin(isph_ctx* c, int alt) {
  const int nb = c->ncell * NCLS + 1;
  LAUNCH(k_fill_i32, nb, c->D.bincnt, nb, 0);
  LAUNCH(k_rebin_classify, c->n_hi, c->D, alt);
  LAUNCH(k_rebin_rank, c->n_hi, c->D);
  int rc = scan_bins(c, nb);
  if (rc) return rc;
  LAUNCH(k_rebin_scatter, c->n_hi, c->D);
  swap_state(c->D);
  LAUNCH(k_rebin_finish, c->ncell + 1, c->D, c->ncell);
  c->assembled = false;
  c->have_nbr = false;
  return 0;
}

extern "C" int isph_cellshift(isph_ctx* c) { return c ? rebin(c, 0) : ISPH_EARG; }
extern "C" int isph_cellshiftalt(isph_ctx* c) { return c ? rebin(c, 1) : ISPH_EARG; }
extern "C" int isph_resetid(isph_ctx* c) {
  // matid is the sorted index: nothing to do on the device (isph.f90:88-113)
  return c ? 0 : ISPH_EARG;
}

// the neighbour search in the context's tuned launch shape
static int launch_neighbour(isph_ctx* c) {
  const int sh = c->tshape[TK_NEIGHBOUR];
  if (sh == TS_TILE4 || sh == TS_TILE8) {
    const int rows = sh == TS_TILE4 ? 4 : 8;
    const dim3 grid(std::max(1, ntile_count(rows, c->D.P.cellx, c->D.P.celly)));
    if (sh == TS_TILE4)
      hipLaunchKernelGGL(k_neighbour_tile4, grid, dim3(128), 0, c->stream, c->D);
    else
      hipLaunchKernelGGL(k_neighbour_tile8, grid, dim3(256), 0, c->stream, c->D);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_err(c, "launch of the tiled neighbour search failed: %s", hipGetErrorString(e));
      return ISPH_EHIP;
    }
    return 0;
  }
  TLAUNCH(TK_NEIGHBOUR, k_neighbour, c->n_hi, c->D);
  return 0;
}

static int neighbour_nocheck(isph_ctx* c) {
  const int rc = launch_neighbour(c);
  if (rc) return rc;
  c->have_nbr = true;
  return 0;
}

extern "C" int isph_neighbour(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  int rc = neighbour_nocheck(c);
  if (rc) return rc;
  return check_err_flags(c);
}

extern "C" int isph_effpor(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  TLAUNCH(TK_EFFPOR, k_effpor, c->n_hi, c->D);
  return 0;
}
extern "C" int isph_eddyvis(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  if (!c->have_nbr) { set_err(c, "eddyvis before neighbour"); return ISPH_ESTATE; }
  TLAUNCH(TK_EDDYVIS, k_eddyvis, c->n_hi, c->D);
  return 0;
}
extern "C" int isph_int_vel(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  if (!c->have_nbr) { set_err(c, "int_vel before neighbour"); return ISPH_ESTATE; }
  TLAUNCH(TK_VARTS, k_int_vel_varts, c->n_hi, c->D);
  TLAUNCH(TK_INTVEL, k_int_vel, c->n_hi, c->D);
  return 0;
}

__global__ void k_zero_nnz(Dev D) {
  if (blockIdx.x == 0 && threadIdx.x == 0) D.sc->nnz = 0;
}

static int assemble_nocheck(isph_ctx* c) {
  hipLaunchKernelGGL(k_zero_nnz, dim3(1), dim3(64), 0, c->stream, c->D);
  TLAUNCH(TK_ASSEMBLE, k_ppe_assemble, c->n_hi, c->D);
  c->assembled = true;
  return 0;
}

extern "C" int isph_ppe_assemble(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  if (!c->have_nbr) { set_err(c, "ppe assembly before neighbour"); return ISPH_ESTATE; }
  int rc = assemble_nocheck(c);
  if (rc) return rc;
  return check_err_flags(c);
}

// ---------------------------------------------------------------------
// precision-form solve for a set of scenarios (one context = a batch of one)
// ---------------------------------------------------------------------
struct BatchHost {
  std::vector<isph_ctx*> ctx;
  int S = 0;
  Dev* d_devs = nullptr;
  MgDev* d_mgs = nullptr;
  int* d_pre = nullptr;
  DevScalars* d_sc = nullptr;
  std::vector<int> pre;          // BK_NKIND x (S+1)
  std::vector<Dev> devs;
  std::vector<MgDev> mgs;
  std::vector<DevScalars> sc;
  int max_nmb = 0, max_nlev = 0;
  int co_lds = 0;                // LDS bytes of the coarse launch (largest scenario's)
};

static void bh_free(BatchHost* b) {
  if (!b) return;
  (void)hipFree(b->d_devs);
  (void)hipFree(b->d_mgs);
  (void)hipFree(b->d_pre);
  (void)hipFree(b->d_sc);
  delete b;
}

// device tables for the scenarios of ctxs (all precision-form contexts)
static int bh_make(isph_ctx* const* ctxs, int S, BatchHost** out) {
  isph_ctx* c = ctxs[0];
  BatchHost* b = new BatchHost();
  b->ctx.assign(ctxs, ctxs + S);
  b->S = S;
  b->pre.assign((size_t)BK_NKIND * (S + 1), 0);
  b->devs.resize(S);
  b->mgs.resize(S);
  b->sc.resize(S);
  hipError_t e = hipMalloc((void**)&b->d_devs, sizeof(Dev) * S);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_mgs, sizeof(MgDev) * S);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_pre, sizeof(int) * b->pre.size());
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_sc, sizeof(DevScalars) * S);
  if (e != hipSuccess) {
    set_err(c, "batch tables: %s", hipGetErrorString(e));
    bh_free(b);
    return ISPH_EHIP;
  }
  // cell-level prefixes are fixed for the life of the contexts
  for (int s = 0; s < S; s++) {
    const MgDev& G = ctxs[s]->G;
    b->max_nmb = std::max(b->max_nmb, G.nmb);
    b->max_nlev = std::max(b->max_nlev, G.nlev);
    b->co_lds = std::max(b->co_lds, ctxs[s]->co_lds);
  }
  for (int s = 0; s < S; s++) {
    const MgDev& G = ctxs[s]->G;
    auto put = [&](int kind, int blocks) {
      b->pre[(size_t)kind * (S + 1) + s + 1] = b->pre[(size_t)kind * (S + 1) + s] + blocks;
    };
    put(BK_CELL1, grid_of(G.lev[1].n));
    for (int l = 1; l < MG_MAX_LEVELS; l++) {
      put(BK_DOWN + l - 1, (l <= G.nmb) ? grid_of(4 * G.lev[l + 1].n) : 0);
      put(BK_UP + l - 1, (l <= G.nmb) ? grid_of((l == 1 ? 4 : 1) * G.lev[l].n) : 0);
      put(BK_COARSEN + l - 1, (l < G.nlev) ? grid_of(G.lev[l + 1].n) : 0);
    }
  }
  *out = b;
  return 0;
}

// refresh the scenario tables (state pointers swap at every rebin) and the
// row prefix (the live count can shrink)
static int bh_refresh(BatchHost* b) {
  isph_ctx* c = b->ctx[0];
  const int S = b->S;
  for (int s = 0; s < S; s++) {
    b->devs[s] = b->ctx[s]->D;
    b->mgs[s] = b->ctx[s]->G;
    b->pre[(size_t)BK_ROWS * (S + 1) + s + 1] =
        b->pre[(size_t)BK_ROWS * (S + 1) + s] + b->ctx[s]->D.nblk;
  }
  HIPCHK(c, hipMemcpyAsync(b->d_devs, b->devs.data(), sizeof(Dev) * S, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(b->d_mgs, b->mgs.data(), sizeof(MgDev) * S, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(b->d_pre, b->pre.data(), sizeof(int) * b->pre.size(),
                           hipMemcpyHostToDevice, c->stream));
  return 0;
}

static int bh_gather(BatchHost* b) {
  isph_ctx* c = b->ctx[0];
  BatchDev B{b->S, b->d_devs, b->d_mgs, b->d_pre, 0};
  hipLaunchKernelGGL(kb_gather_scalars, dim3((b->S + 63) / 64), dim3(64), 0, c->stream, B, b->d_sc);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(b->sc.data(), b->d_sc, sizeof(DevScalars) * b->S,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

#define BLAUNCH(kern, grid, block, ...) BLAUNCH_LDS(kern, grid, block, 0, __VA_ARGS__)
#define BLAUNCH_LDS(kern, grid, block, lds, ...)                                      \
  do {                                                                                \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, c->stream, __VA_ARGS__);   \
    hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess) {                                                           \
      set_err(c, "launch of %s failed: %s", #kern, hipGetErrorString(e_));           \
      return ISPH_EHIP;                                                               \
    }                                                                                 \
  } while (0)

// A timed batched launch (events attached to the dispatch, as TIMED_LAUNCH)
#define TIMED_BLAUNCH_LDS(tidx, kern, grid, block, lds, ...)                          \
  do {                                                                                \
    if (c->prof) {                                                                    \
      std::pair<hipEvent_t, hipEvent_t> ev_ = timer_pair(c->timers[tidx]);            \
      hipExtLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, c->stream, ev_.first, \
                            ev_.second, 0, __VA_ARGS__);                              \
      hipError_t e_ = hipGetLastError();                                              \
      if (e_ != hipSuccess) {                                                         \
        set_err(c, "launch of %s failed: %s", #kern, hipGetErrorString(e_));         \
        return ISPH_EHIP;                                                             \
      }                                                                               \
      c->timers[tidx].pending.push_back(ev_);                                         \
    } else {                                                                          \
      BLAUNCH_LDS(kern, grid, block, lds, __VA_ARGS__);                               \
    }                                                                                 \
  } while (0)
#define TIMED_BLAUNCH(tidx, kern, grid, ...) \
  TIMED_BLAUNCH_LDS(tidx, kern, grid, 256, 0, __VA_ARGS__)

// one V-cycle for every scenario: WHICH 0 gives r32 = M(pp32), 1 sv32 = M(sv);
// OFF: level-0 rows read the packed 16-bit column offsets
template <int WHICH, bool OFF>
static int bh_vcycle(BatchHost* b, const BatchDev& B, int grow) {
  constexpr int V = WHICH | (OFF ? 2 : 0);
  isph_ctx* c = b->ctx[0];
  const int S = b->S;
  auto tot = [&](int kind) { return b->pre[(size_t)kind * (S + 1) + S]; };
  TIMED_BLAUNCH(5, kb_mg_pre0<V>, grow, B);
  TIMED_BLAUNCH(10, kb_mg_restrict0, tot(BK_CELL1), B);
  for (int l = 1; l <= b->max_nmb; l++)
    if (tot(BK_DOWN + l - 1) > 0) TIMED_BLAUNCH(7, kb_mg_down, tot(BK_DOWN + l - 1), B, l);
  TIMED_BLAUNCH_LDS(4, kb_mg_coarse, S, MG_CO_THREADS, b->co_lds, B);
  for (int l = b->max_nmb; l >= 1; l--)
    if (tot(BK_UP + l - 1) > 0) TIMED_BLAUNCH(8, kb_mg_up, tot(BK_UP + l - 1), B, l);
  TIMED_BLAUNCH(6, kb_mg_post0<V>, grow, B);
  return 0;
}

// chunk iterations of the preconditioned BiCGStab for every scenario; seq
// numbers the iterations launched in this solve (the finalize that stops a
// scenario records it, so that iteration's x and r update alone still runs)
template <bool OFF>
static int bh_iterate(BatchHost* b, const BatchDev& B, int grow, int chunk, int& seq) {
  isph_ctx* c = b->ctx[0];
  const int S = b->S;
  int rc = 0;
  for (int k = 0; k < chunk; k++, seq++) {
    rc = bh_vcycle<0, OFF>(b, B, grow);
    if (rc) return rc;
    TIMED_BLAUNCH(2, kb_pb_1<OFF>, grow, B);
    TIMED_BLAUNCH(9, kb_pfinalize, S, B, (int)PFIN_ALPHA, 0, seq);
    BLAUNCH(kb_pb_s, grow, 256, B);
    rc = bh_vcycle<1, OFF>(b, B, grow);
    if (rc) return rc;
    TIMED_BLAUNCH(3, kb_pb_2<OFF>, grow, B);
    TIMED_BLAUNCH(9, kb_pfinalize, S, B, (int)PFIN_OMEGA, 0, seq);
    BLAUNCH(kb_pb_34, grow, 256, B, seq);
  }
  return 0;
}

// Precision-form solve of every scenario of b: cell-multigrid right-
// preconditioned BiCGStab with a V-cycle in FP16 (level-0 matrix) and FP32
// (vectors) and FP32 search directions; FP64 outer products A ph and A sh,
// solution, residual and dots; every solve starts from x = 0.  A recursive-residual convergence claim
// is checked against b - A64 x after the chunk in which it was made; when the
// check fails the iteration continues from the FP64 residual.  The host only
// polls whether every scenario has finished; stats[s] gets scenario s's.
static int bh_solve(BatchHost* b, isph_solve_stats* stats) {
  isph_ctx* c = b->ctx[0];
  const int S = b->S;
  int rc = bh_refresh(b);
  if (rc) return rc;
  auto tot = [&](int kind) { return b->pre[(size_t)kind * (S + 1) + S]; };
  const BatchDev B{S, b->d_devs, b->d_mgs, b->d_pre, tot(BK_ROWS)};
  c->timers[7].per_iter = c->timers[8].per_iter = 2 * b->max_nmb;  // mg_down, mg_up
  const int grow = 8 * ((tot(BK_ROWS) + 7) / 8);
  // multigrid hierarchy from the assembled rows
  BLAUNCH(kb_mg_setup0, grow, 256, B);
  BLAUNCH(kb_mg_stencil1, tot(BK_CELL1), 256, B);
  for (int l = 1; l < b->max_nlev; l++)
    if (tot(BK_COARSEN + l - 1) > 0) BLAUNCH(kb_mg_coarsen, tot(BK_COARSEN + l - 1), 256, B, l);
  BLAUNCH(kb_mg_dense, S, MG_CO_THREADS, B);
  // ||b||, then the FP64 start residual
  BLAUNCH(kb_pb_bnorm, grow, 256, B);
  BLAUNCH(kb_pfinalize, S, 256, B, (int)PFIN_BNORM, 0, 0);
  BLAUNCH(kb_pb_resid, grow, 256, B, 1);
  BLAUNCH(kb_pfinalize, S, 256, B, (int)PFIN_RESID, 1, 0);
  rc = bh_gather(b);
  if (rc) return rc;
  // error bits of the phases before the solve (neighbour capacity, PPE
  // diagonal) end the step, as the per-scenario check does
  for (int s = 0; s < S; s++) {
    if (b->sc[s].err) {
      rc = check_err_flags(b->ctx[s]);
      if (rc) return rc;
    }
  }
  // packed 16-bit column offsets unless some scenario's rows need int32
  bool off16 = true;
  for (int s = 0; s < S; s++) off16 = off16 && !b->sc[s].far;
  std::vector<int> it_prev(S), rs_prev(S), first_stop(S, -1);
  auto all_fin = [&]() {
    for (int s = 0; s < S; s++)
      if (!b->sc[s].fin) return false;
    return true;
  };
  int max_iter = 0, max_outer = 0;
  for (int s = 0; s < S; s++) {
    max_iter = std::max(max_iter, b->ctx[s]->P.max_iter);
    max_outer = std::max(max_outer, b->ctx[s]->P.max_outer);
    it_prev[s] = b->sc[s].iters;
    rs_prev[s] = b->sc[s].restarts;
  }
  // every chunk either advances an unfinished scenario by at least one
  // iteration or ends it at the check, so this bound is never the exit
  const int64_t chunk_cap = 4 + (int64_t)(max_iter + 1) * (max_outer + 2);
  // Chunk sizes: the iterations the scenarios needed last time to their first
  // stop (and after it), then 2, 4, 8, 16 until they stop.  A chunk ends in a
  // check; iterations launched after a stop exit at once.
  auto predicted = [&](bool after) {
    int p = 0;
    for (int s = 0; s < S; s++)
      if (!b->sc[s].fin) p = std::max(p, after ? b->ctx[s]->pred_after : b->ctx[s]->pred_first);
    return p;
  };
  int chunk = 4, seq = 0;
  bool guessed = false;  // the chunk just run was sized from the last solve
  if (const int p = predicted(false); p > 0) {
    chunk = p;
    guessed = true;
  }
  for (int64_t nchunks = 0; !all_fin(); nchunks++) {
    if (nchunks > chunk_cap) {
      set_err(c, "precision solve: scenarios did not finish");
      return ISPH_ENOCONV;
    }
    rc = off16 ? bh_iterate<true>(b, B, grow, chunk, seq)
               : bh_iterate<false>(b, B, grow, chunk, seq);
    if (rc) return rc;
    // FP64 check of every scenario whose iteration stopped in this chunk
    BLAUNCH(kb_pb_resid, grow, 256, B, 2);
    BLAUNCH(kb_pfinalize, S, 256, B, (int)PFIN_RESID, 2, 0);
    rc = bh_gather(b);
    if (rc) return rc;
    // the timed launches that did work: the chunk's leading iterations
    int worked = 0, restarted = 0;
    for (int s = 0; s < S; s++) {
      worked = std::max(worked, b->sc[s].iters - it_prev[s]);
      if (b->sc[s].restarts != rs_prev[s]) {
        restarted = 1;
        if (first_stop[s] < 0) first_stop[s] = b->sc[s].iters;
      }
      it_prev[s] = b->sc[s].iters;
      rs_prev[s] = b->sc[s].restarts;
    }
    if (c->prof) timers_collect(c, worked);
    if (restarted) {
      const int p = predicted(true);
      guessed = p > 0;
      chunk = guessed ? p : 4;
    } else {
      chunk = guessed ? 2 : std::min(16, 2 * chunk);
      guessed = false;
    }
  }
  for (int s = 0; s < S; s++) {
    const int tot_it = b->sc[s].iters;
    b->ctx[s]->pred_first = (first_stop[s] >= 0) ? first_stop[s] : tot_it;
    b->ctx[s]->pred_after = (first_stop[s] >= 0) ? tot_it - first_stop[s] : 0;
  }
  BLAUNCH(kb_copy_sol_to_p, grow, 256, B);
  for (int s = 0; s < S; s++) {
    const DevScalars& d = b->sc[s];
    isph_solve_stats& st = stats[s];
    st.iterations = d.iters;
    st.outer = d.restarts;
    st.converged = d.conv;
    st.breakdown = d.brk;
    st.rel_residual = d.rel;
    st.bnorm = d.bnorm;
    st.nnz = (double)d.nnz;
  }
  return 0;
}

// run BiCGStab chunks until the device reports done; on a breakdown below
// the iteration cap restart from the true residual, at most max_outer times
static int bicg_loop(isph_ctx* c, int* iters_out, int* brk_out, int* conv_out,
                     double* rel_out, int* restarts_out) {
  const int n = c->n_hi;
  int chunk = 4, restarts = 0, prev = 0;
  DevScalars s;
  for (int64_t it = 0;; ) {
    for (int k = 0; k < chunk; k++) {
      TIMED_LAUNCH(0, k_bicg_1, n, c->D);
      FINALIZE(FIN_ALPHA, n);
      TIMED_LAUNCH(1, k_bicg_2, n, c->D);
      FINALIZE(FIN_OMEGA, n);
      RLAUNCH(k_bicg_3, n, c->D);
      FINALIZE(FIN_BETA, n);
      RLAUNCH(k_bicg_4, n, c->D);
    }
    it += chunk;
    int rc = read_scalars(c, &s);
    if (rc) return rc;
    // the timed launches that did work: the chunk's leading iterations
    if (c->prof) timers_collect(c, s.iters - prev);
    prev = s.iters;
    if (s.done) {
      if (s.brk && !s.conv && s.iters < c->P.max_iter && restarts < c->P.max_outer) {
        restarts++;
        RLAUNCH(k_bicg_init, n, c->D);
        FINALIZE(FIN_RESTART, n);
        rc = read_scalars(c, &s);
        if (rc) return rc;
        if (s.done) break;
        chunk = 4;
        continue;
      }
      break;
    }
    if (chunk < 32) chunk *= 2;
    if (it > (int64_t)(c->P.max_iter + 64) * (c->P.max_outer + 1)) break;
  }
  *iters_out = s.iters;
  *brk_out = s.brk;
  *conv_out = s.conv;
  *rel_out = s.rel;
  *restarts_out = restarts;
  return 0;
}

static int ppe_solve_impl(isph_ctx* c, isph_solve_stats* stats) {
  if (!c->assembled) {
    set_err(c, "ppe solve before assembly");
    return ISPH_ESTATE;
  }
  const int n = c->n_hi;
  isph_solve_stats st{};
  if (c->P.precision != ISPH_PPE_FP64) {
    if (!c->self_batch) {
      int rc = bh_make(&c, 1, &c->self_batch);
      if (rc) return rc;
    }
    int rc = bh_solve(c->self_batch, &st);
    if (rc) return rc;
  } else {
    {
      long long nnz = 0;
      HIPCHK(c, hipMemcpyAsync(&nnz, &c->D.sc->nnz, sizeof nnz, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      st.nnz = (double)nnz;
    }
    // ||b||
    RLAUNCH(k_dot, n, c->D, c->D.b, c->D.b);
    FINALIZE_OUT(FIN_DOT, n, c->d_bn2);
    double bn2 = 0;
    HIPCHK(c, hipMemcpyAsync(&bn2, c->d_bn2, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double bnorm = std::sqrt(bn2);
    st.bnorm = bnorm;
    HIPCHK(c, hipMemcpy(&c->D.sc->bnorm, &bnorm, 8, hipMemcpyHostToDevice));
    if (bnorm == 0.0 || !std::isfinite(bnorm)) {
      if (bnorm == 0.0) {
        HIPCHK(c, hipMemsetAsync(c->D.sol, 0, 8 * (size_t)c->cap, c->stream));
        st.converged = 1;
      }
      st.rel_residual = (bnorm == 0.0) ? 0.0 : bnorm;
    } else {
      RLAUNCH(k_bicg_init, n, c->D);
      FINALIZE(FIN_INIT, n);
      DevScalars s;
      int rc = read_scalars(c, &s);
      if (rc) return rc;
      int it = 0, brk = 0, conv = s.conv, restarts = 0;
      double rel = s.rel;
      if (!s.done) {
        rc = bicg_loop(c, &it, &brk, &conv, &rel, &restarts);
        if (rc) return rc;
      }
      st.iterations = it;
      st.outer = restarts;
      st.breakdown = brk;
      st.converged = conv;
      st.rel_residual = rel;
    }
    LAUNCH(k_copy_sol_to_p, n, c->D);
  }
  c->last = st;
  if (stats) *stats = st;
  if (!std::isfinite(st.rel_residual) || !std::isfinite(st.bnorm)) {
    set_err(c, "non-finite PPE residual");
    return ISPH_ENONFINITE;
  }
  if (!st.converged) {
    set_err(c, "PPE solve did not converge: %d iterations, rel residual %g", st.iterations,
            st.rel_residual);
    return ISPH_ENOCONV;
  }
  return 0;
}

extern "C" int isph_ppe_solve(isph_ctx* c, isph_solve_stats* stats) {
  if (!c) return ISPH_EARG;
  return ppe_solve_impl(c, stats);
}

extern "C" int isph_ppesolve(isph_ctx* c, isph_solve_stats* stats) {
  int rc = isph_ppe_assemble(c);
  if (rc) return rc;
  return isph_ppe_solve(c, stats);
}

extern "C" int isph_comp_vel(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  TLAUNCH(TK_COMPVEL, k_comp_vel, c->n_hi, c->D);
  LAUNCH(k_cell_max, c->ncell, c->D, c->ncell, 0);
  return 0;
}
extern "C" int isph_comp_pos(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  LAUNCH(k_comp_pos, c->n_hi, c->D);
  return 0;
}
extern "C" int isph_freesurf(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  LAUNCH(k_freesurf, c->n_hi, c->D);
  return 0;
}
extern "C" int isph_opt2_shift(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  TLAUNCH(TK_SHIFT, k_opt2_shift, c->n_hi, c->D);
  LAUNCH(k_shift_apply, c->n_hi, c->D);
  return 0;
}
extern "C" int isph_densityupdate(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  LAUNCH(k_densityupdate, c->n_hi, c->D);
  return 0;
}

extern "C" int isph_timestep(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  hipLaunchKernelGGL(k_timestep, dim3(1), dim3(256), 0, c->stream, c->D);
  HIPCHK(c, hipGetLastError());
  return 0;
}
extern "C" int isph_scalart_anisotropic(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  if (!c->have_nbr) { set_err(c, "scalar transport before neighbour"); return ISPH_ESTATE; }
  TLAUNCH(TK_SCALART, k_scalart, c->n_hi, c->D);
  return 0;
}

extern "C" int isph_scalarupdate(isph_ctx* c, double stime) {
  if (!c || !(stime >= 0.0)) return ISPH_EARG;
  LAUNCH(k_scalarupdate, c->n_hi, c->D, stime);
  return 0;
}

extern "C" int isph_init_state(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  int rc = isph_neighbour(c);
  if (rc) return rc;
  TLAUNCH(TK_EFFPOR, k_effpor, c->n_hi, c->D);
  TLAUNCH(TK_EDDYVIS, k_eddyvis, c->n_hi, c->D);
  LAUNCH(k_cell_max, c->ncell, c->D, c->ncell, 1);
  rc = isph_timestep(c);
  if (rc) return rc;
  return sync(c);
}

__global__ void k_advance_t(Dev D) {
  if (blockIdx.x == 0 && threadIdx.x == 0) D.sc->t = D.sc->t + D.sc->dt;
}

// ---------------------------------------------------------------------
// inflow/outflow buffer zones (internal.f90:447-586, main.f90:217-235)
// ---------------------------------------------------------------------
extern "C" int isph_set_buffer(isph_ctx* c, int32_t zone, int64_t nt, const double* x,
                               const double* y, const double* u, const double* v,
                               const double* p, double mass, double rho, double vol0,
                               double cval, double rv, double interval, double time_shift) {
  if (!c || zone < 1 || zone > 2 || nt < 0 || nt > c->cap) return ISPH_EARG;
  if (nt > 0 && (!x || !y || !u || !v || !p)) return ISPH_EARG;
  HostZone& Z = c->zones[zone - 1];
  HIPCHK(c, hipStreamSynchronize(c->stream));  // the zone's arrays may be in use
  std::vector<int32_t> tcell(nt), trank(nt), tcnt(c->ncell, 0);
  std::vector<uint8_t> entry(c->ncell, 0);
  for (int64_t t = 0; t < nt; t++) {
    const int k = locate_cell(c->P, x[t], y[t]);
    if (k < 0) {
      set_err(c, "buffer template particle %lld at (%g, %g) is outside the interior cells",
              (long long)t, x[t], y[t]);
      return ISPH_EARG;
    }
    tcell[t] = k;
    trank[t] = tcnt[k]++;
    entry[k] = 1;
  }
  void* zp[] = {Z.tx, Z.ty, Z.tu, Z.tv, Z.tp, Z.tcell, Z.trank, Z.tcnt, Z.entry};
  for (void* q : zp) (void)hipFree(q);
  Z = HostZone();
  const size_t m = (size_t)std::max<int64_t>(nt, 1);
  hipError_t e = hipMalloc((void**)&Z.tx, 8 * m);
  if (e == hipSuccess) e = hipMalloc((void**)&Z.ty, 8 * m);
  if (e == hipSuccess) e = hipMalloc((void**)&Z.tu, 8 * m);
  if (e == hipSuccess) e = hipMalloc((void**)&Z.tv, 8 * m);
  if (e == hipSuccess) e = hipMalloc((void**)&Z.tp, 8 * m);
  if (e == hipSuccess) e = hipMalloc((void**)&Z.tcell, 4 * m);
  if (e == hipSuccess) e = hipMalloc((void**)&Z.trank, 4 * m);
  if (e == hipSuccess) e = hipMalloc((void**)&Z.tcnt, 4 * (size_t)c->ncell);
  if (e == hipSuccess) e = hipMalloc((void**)&Z.entry, (size_t)c->ncell);
  if (e != hipSuccess) {
    set_err(c, "buffer zone allocation: %s", hipGetErrorString(e));
    return ISPH_EHIP;
  }
  if (nt > 0) {
    HIPCHK(c, hipMemcpy(Z.tx, x, 8 * nt, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(Z.ty, y, 8 * nt, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(Z.tu, u, 8 * nt, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(Z.tv, v, 8 * nt, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(Z.tp, p, 8 * nt, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(Z.tcell, tcell.data(), 4 * nt, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(Z.trank, trank.data(), 4 * nt, hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipMemcpy(Z.tcnt, tcnt.data(), 4 * (size_t)c->ncell, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(Z.entry, entry.data(), (size_t)c->ncell, hipMemcpyHostToDevice));
  Z.nt = (int)nt;
  Z.mass = mass; Z.rho = rho; Z.vol0 = vol0; Z.c = cval; Z.rv = rv;
  Z.interval = interval; Z.time_shift = time_shift;
  Z.iter = 1;
  Z.set = true;
  return 0;
}

extern "C" int isph_buffer_update(isph_ctx* c, int32_t zone) {
  if (!c || zone < 1 || zone > 2) return ISPH_EARG;
  HostZone& Zh = c->zones[zone - 1];
  if (!Zh.set) {
    set_err(c, "buffer zone %d is not set", zone);
    return ISPH_ESTATE;
  }
  if ((int64_t)c->n_hi + Zh.nt > c->cap) {
    set_err(c, "buffer insertion could exceed capacity %lld", (long long)c->cap);
    return ISPH_ECAP;
  }
  BufZone Z{zone, Zh.nt, Zh.tx, Zh.ty, Zh.tu, Zh.tv, Zh.tp, Zh.tcell, Zh.trank, Zh.tcnt,
            Zh.entry, Zh.mass, Zh.rho, Zh.vol0, Zh.c, Zh.rv};
  const int nc = c->ncell;
  LAUNCH(k_fill_i32, 2 * nc + 1, c->D.bincnt, 2 * nc + 1, 0);
  LAUNCH(k_buf_rank, nc, c->D, Z, nc);
  int rc = scan_bins(c, nc + 1);
  if (rc) return rc;
  LAUNCH(k_buf_scatter_kept, c->n_hi, c->D);
  if (Zh.nt > 0) LAUNCH(k_buf_scatter_new, Zh.nt, c->D, Z, nc);
  swap_state(c->D);
  const int n_new_hi = c->n_hi + Zh.nt;
  LAUNCH(k_buf_finish, std::max(nc + 1, n_new_hi), c->D, nc, Zh.nt);
  hipLaunchKernelGGL(k_buf_set_n, dim3(1), dim3(64), 0, c->stream, c->D, nc);
  HIPCHK(c, hipGetLastError());
  set_nhi(c, n_new_hi);
  c->assembled = false;
  c->have_nbr = false;
  return 0;
}

// ---------------------------------------------------------------------
// weakly compressible variant (wcsph.f90:13-154)
// ---------------------------------------------------------------------
extern "C" int isph_set_wcsph(isph_ctx* c, double c0, double gamma, double delta,
                              double swl, int32_t in_step) {
  if (!c || !(c0 > 0.0) || !(gamma > 0.0)) return ISPH_EARG;
  if (c->in_batch) {
    set_err(c, "the weakly compressible form cannot be set while the context belongs to a batch");
    return ISPH_ESTATE;
  }
  c->D.P.c0 = c0;
  c->D.P.gam = gamma;
  c->D.P.delt = delta;
  c->D.P.swl = swl;
  c->wcsph = in_step != 0;
  return 0;
}
extern "C" int isph_presstoden(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  LAUNCH(k_presstoden, c->n_hi, c->D);
  return 0;
}
extern "C" int isph_dentopress(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  LAUNCH(k_dentopress, c->n_hi, c->D);
  return 0;
}
extern "C" int isph_calmasscons(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  if (!c->have_nbr) { set_err(c, "calmasscons before neighbour"); return ISPH_ESTATE; }
  LAUNCH(k_calmasscons, c->n_hi, c->D);
  return 0;
}
extern "C" int isph_updateden(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  LAUNCH(k_updateden, c->n_hi, c->D);
  return 0;
}

// ---------------------------------------------------------------------
// piston wavemaker (no counterpart in the reference)
// ---------------------------------------------------------------------
extern "C" int isph_wavemaker_motion(double stroke, double period, double ramp, double t0,
                                     double t, double* x, double* u) {
  if (!x || !u || !(period > 0.0) || !(stroke >= 0.0) || !(ramp >= 0.0)) return ISPH_EARG;
  const double tau = t - t0;
  if (!(tau > 0.0)) {
    *x = 0.0;
    *u = 0.0;
    return 0;
  }
  const double a = 0.5 * stroke;
  const double w = (2.0 * M_PI) / period;
  double r = 1.0, rp = 0.0;
  if (ramp > 0.0 && tau < ramp) {
    r = tau / ramp;
    rp = 1.0 / ramp;
  }
  const double sn = std::sin(w * tau), cs = std::cos(w * tau);
  *x = (a * r) * sn;
  *u = a * (((r * w) * cs) + (rp * sn));
  return 0;
}

// the paddle at the motion's value for time t
static int paddle_to(isph_ctx* c, double t) {
  double X, U;
  if (isph_wavemaker_motion(c->wv_stroke, c->wv_period, c->wv_ramp, c->wv_t0, t, &X, &U))
    return ISPH_EARG;
  LAUNCH(k_paddle, c->n_hi, c->D, X, U);
  return 0;
}

extern "C" int isph_set_wavemaker(isph_ctx* c, double stroke, double period, double ramp,
                                  double t0) {
  if (!c) return ISPH_EARG;
  double X, U;
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  if (isph_wavemaker_motion(stroke, period, ramp, t0, s.t, &X, &U)) {
    set_err(c, "wavemaker needs stroke >= 0, period > 0 and ramp >= 0 (got %g, %g, %g)",
            stroke, period, ramp);
    return ISPH_EARG;
  }
  set_nhi(c, s.n);
  LAUNCH(k_paddle_rest, c->n_hi, c->D, X);
  c->wv_stroke = stroke;
  c->wv_period = period;
  c->wv_ramp = ramp;
  c->wv_t0 = t0;
  c->wave = true;
  return 0;
}

extern "C" int isph_wavemaker(isph_ctx* c) {
  if (!c) return ISPH_EARG;
  if (!c->wave) {
    set_err(c, "wavemaker is not set");
    return ISPH_ESTATE;
  }
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  set_nhi(c, s.n);
  return paddle_to(c, s.t);
}

// the insertion clock of main.f90:217-235 (told = t - dt)
static int buffer_clock(isph_ctx* c, double told, double t) {
  for (int z = 0; z < 2; z++) {
    HostZone& Z = c->zones[z];
    if (!Z.set || !(Z.interval > 0.0)) continue;
    if (((told + Z.time_shift) < Z.iter * Z.interval) &&
        ((t + Z.time_shift) >= Z.iter * Z.interval)) {
      int rc = isph_buffer_update(c, z + 1);
      if (rc) return rc;
      Z.iter = Z.iter + 1;
    }
  }
  return 0;
}

// one production-loop iteration, main.f90:199-261, in three parts: up to the
// PPE assembly, the solve, and the rest
static int step_pre(isph_ctx* c) {
  DevScalars s;
  int rc = read_scalars(c, &s);  // solsteps from the previous timestep
  if (rc) return rc;
  // error bits the previous step's last phases set (its final check is left
  // to this read when another step follows)
  if (s.err && (rc = check_err_flags(c))) return rc;
  set_nhi(c, s.n);
  if (s.solsteps < 1 || s.solsteps > ISPH_MAX_SOLSTEPS) {
    set_err(c, "scalar sub-step count %d outside [1, %d]", s.solsteps, ISPH_MAX_SOLSTEPS);
    return ISPH_ENONFINITE;
  }
  hipLaunchKernelGGL(k_advance_t, dim3(1), dim3(64), 0, c->stream, c->D);
  for (int q = 0; q < s.solsteps; q++) {
    TLAUNCH(TK_SCALART, k_scalart, c->n_hi, c->D);
    LAUNCH(k_scalarupdate, c->n_hi, c->D, -1.0);
  }
  rc = buffer_clock(c, s.t, s.t + s.dt);  // the device's t + dt, the same double sum
  if (rc) return rc;
  LAUNCH(k_projection, c->n_hi, c->D);
  if (c->wave) {
    rc = paddle_to(c, s.t + s.dt);  // the device's t + dt, the same double sum
    if (rc) return rc;
  }
  rc = rebin(c, 0);
  if (rc) return rc;
  rc = neighbour_nocheck(c);
  if (rc) return rc;
  TLAUNCH(TK_EFFPOR, k_effpor, c->n_hi, c->D);
  TLAUNCH(TK_EDDYVIS, k_eddyvis, c->n_hi, c->D);
  TLAUNCH(TK_VARTS, k_int_vel_varts, c->n_hi, c->D);
  TLAUNCH(TK_INTVEL, k_int_vel, c->n_hi, c->D);
  if (c->wcsph) {
    // continuity and the Tait law in place of the pressure Poisson solve
    LAUNCH(k_calmasscons, c->n_hi, c->D);
    LAUNCH(k_updateden, c->n_hi, c->D);
    LAUNCH(k_dentopress, c->n_hi, c->D);
    return check_err_flags(c);
  }
  rc = assemble_nocheck(c);
  if (rc) return rc;
  // the precision-form solve reads the error bits with its first scalars and
  // ends the step on them, so the GPU does not wait here for the host
  if (c->P.precision != ISPH_PPE_FP64) return 0;
  return check_err_flags(c);
}

// check: read the error bits now (else the next step_pre's scalar read does)
static int step_post(isph_ctx* c, bool check) {
  TLAUNCH(TK_COMPVEL, k_comp_vel, c->n_hi, c->D);
  LAUNCH(k_cell_max, c->ncell, c->D, c->ncell, 0);
  LAUNCH(k_comp_pos, c->n_hi, c->D);
  int rc = rebin(c, 1);
  if (rc) return rc;
  rc = neighbour_nocheck(c);
  if (rc) return rc;
  LAUNCH(k_freesurf, c->n_hi, c->D);
  TLAUNCH(TK_EFFPOR, k_effpor, c->n_hi, c->D);
  TLAUNCH(TK_SHIFT, k_opt2_shift, c->n_hi, c->D);
  LAUNCH(k_shift_apply, c->n_hi, c->D);
  if (!c->wcsph) LAUNCH(k_densityupdate, c->n_hi, c->D);
  rc = isph_timestep(c);
  if (rc) return rc;
  TLAUNCH(TK_EDDYVIS, k_eddyvis, c->n_hi, c->D);
  return check ? check_err_flags(c) : 0;
}

// a solve result that ends the step: errors other than non-convergence (the
// error bits of the phases before the solve, read by the precision form's
// solve, included)
static bool solve_fatal(int rc) { return rc != 0 && rc != ISPH_ENOCONV; }

static int step_once(isph_ctx* c, isph_solve_stats* stats, int* solve_rc, bool last) {
  int rc = step_pre(c);
  if (rc) return rc;
  if (c->wcsph) {
    if (stats) *stats = isph_solve_stats{};
    return step_post(c, last);
  }
  rc = ppe_solve_impl(c, stats);
  if (solve_fatal(rc)) return rc;
  *solve_rc = rc;
  return step_post(c, last);
}

extern "C" int isph_step(isph_ctx* c, int32_t nsteps, isph_solve_stats* stats) {
  if (!c || nsteps < 0) return ISPH_EARG;
  int worst = 0;
  for (int k = 0; k < nsteps; k++) {
    int srcode = 0;
    int rc = step_once(c, stats, &srcode, k == nsteps - 1);
    if (rc) return rc;
    if (srcode) worst = srcode;
  }
  int rc = sync(c);
  if (rc) return rc;
  return worst;
}

// ---------------------------------------------------------------------
// dumps for phase-by-phase parity
// ---------------------------------------------------------------------
extern "C" int isph_dump_cells(isph_ctx* c, int32_t* cell_of, int32_t* cell_start) {
  if (!c) return ISPH_EARG;
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  if (cell_of) HIPCHK(c, hipMemcpy(cell_of, c->D.cell, 4 * (size_t)s.n, hipMemcpyDeviceToHost));
  if (cell_start)
    HIPCHK(c, hipMemcpy(cell_start, c->D.cstart, 4 * (size_t)(c->ncell + 1), hipMemcpyDeviceToHost));
  return 0;
}

template <typename T>
static int dump_ell(isph_ctx* c, const T* src, T* dst, int width, int n) {
  std::vector<T> tmp((size_t)width * c->cap);
  HIPCHK(c, hipMemcpy(tmp.data(), src, sizeof(T) * (size_t)width * c->cap, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++)
    for (int k = 0; k < width; k++) dst[(size_t)i * width + k] = tmp[(size_t)k * c->cap + i];
  return 0;
}

extern "C" int isph_dump_neighbours(isph_ctx* c, int32_t* count, int32_t* list) {
  if (!c) return ISPH_EARG;
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  if (count) HIPCHK(c, hipMemcpy(count, c->D.ncnt, 4 * (size_t)s.n, hipMemcpyDeviceToHost));
  if (list) return dump_ell(c, c->D.nbr, list, NBR_CAP, s.n);
  return 0;
}

extern "C" int isph_dump_props(isph_ctx* c, double* porosity, double* resist, double* gradvx,
                               double* coff, double* nut, double* tke, double* varts,
                               double* cdiff, double* vel, double* us, double* vs, double* xs,
                               double* ys) {
  if (!c) return ISPH_EARG;
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  const size_t n = s.n;
  Dev& D = c->D;
  struct { double* dst; double* src; } f[] = {{porosity, D.por}, {resist, D.resist},
                                              {gradvx, D.gradvx}, {nut, D.nut}, {tke, D.tke},
                                              {varts, D.varts}, {cdiff, D.cdiff}, {vel, D.vel},
                                              {us, D.us}, {vs, D.vs}, {xs, D.xs}, {ys, D.ys}};
  for (auto& e : f)
    if (e.dst) HIPCHK(c, hipMemcpy(e.dst, e.src, 8 * n, hipMemcpyDeviceToHost));
  if (coff) {
    std::vector<double> a(n), b(n), cc(n), d(n);
    HIPCHK(c, hipMemcpy(a.data(), D.cf1, 8 * n, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(b.data(), D.cf2, 8 * n, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(cc.data(), D.cf3, 8 * n, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(d.data(), D.cf4, 8 * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
      coff[4 * i] = a[i]; coff[4 * i + 1] = b[i]; coff[4 * i + 2] = cc[i]; coff[4 * i + 3] = d[i];
    }
  }
  return 0;
}

extern "C" int isph_dump_ppe(isph_ctx* c, int32_t* row_len, int32_t* col, double* val,
                             double* rhs) {
  if (!c) return ISPH_EARG;
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  const int n = s.n;
  if (row_len) HIPCHK(c, hipMemcpy(row_len, c->D.alen, 4 * (size_t)n, hipMemcpyDeviceToHost));
  if (rhs) HIPCHK(c, hipMemcpy(rhs, c->D.b, 8 * (size_t)n, hipMemcpyDeviceToHost));
  if (col && (rc = dump_ell(c, c->D.acol, col, ROW_CAP, n))) return rc;
  if (val && (rc = dump_ell(c, c->D.aval, val, ROW_CAP, n))) return rc;
  return 0;
}

extern "C" int isph_ppe_set_rhs(isph_ctx* c, const double* rhs, const double* x0) {
  if (!c) return ISPH_EARG;
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  if (rhs) HIPCHK(c, hipMemcpy(c->D.b, rhs, 8 * (size_t)s.n, hipMemcpyHostToDevice));
  if (x0) HIPCHK(c, hipMemcpy(c->D.sol, x0, 8 * (size_t)s.n, hipMemcpyHostToDevice));
  return 0;
}

extern "C" int isph_ppe_get_solution(isph_ctx* c, double* x) {
  if (!c || !x) return ISPH_EARG;
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  HIPCHK(c, hipMemcpy(x, c->D.sol, 8 * (size_t)s.n, hipMemcpyDeviceToHost));
  return 0;
}

// Load an explicit system in canonical row form (row-major, ISPH_NBR_CAP+2 wide)
// and mark it assembled: used by the solver known-answer tests.
extern "C" int isph_load_system(isph_ctx* c, int32_t n, const int32_t* len, const int32_t* col,
                                const double* val, const double* rhs, const double* x0) {
  if (!c || n < 0 || n > c->cap) return ISPH_EARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int32_t> L(n), C((size_t)ROW_CAP * c->cap, 0);
  std::vector<double> V((size_t)ROW_CAP * c->cap, 0.0);
  for (int i = 0; i < n; i++) {
    if (len[i] < 0 || len[i] > ROW_CAP) return ISPH_EARG;
    L[i] = len[i];
    for (int k = 0; k < len[i]; k++) {
      const int j = col[(size_t)i * ROW_CAP + k];
      if (j < 0 || j >= n) return ISPH_EARG;
      C[(size_t)k * c->cap + i] = j;
      V[(size_t)k * c->cap + i] = val[(size_t)i * ROW_CAP + k];
    }
  }
  HIPCHK(c, hipMemcpy(c->D.alen, L.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->D.acol, C.data(), 4 * C.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->D.aval, V.data(), 8 * V.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->D.b, rhs, 8 * (size_t)n, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->D.sol, x0, 8 * (size_t)n, hipMemcpyHostToDevice));
  DevScalars s;
  int rc = read_scalars(c, &s);
  if (rc) return rc;
  s.n = n;
  HIPCHK(c, hipMemcpy(c->D.sc, &s, sizeof s, hipMemcpyHostToDevice));
  set_nhi(c, n);
  c->assembled = true;
  return 0;
}

// ---------------------------------------------------------------------
// batched form: independent scenarios advanced together
// ---------------------------------------------------------------------
// Every context of a batch is switched to one shared stream.  When every
// scenario is an ISPH scenario without an insertion clock on its buffer
// zones, a batched step runs each particle phase as one launch over all
// scenarios (isph_phases.h) and the precision-form solves as one set of
// launches (bh_solve): the host reads every scenario's scalars once at the
// start of a step (solsteps, the paddle's time, error bits) and otherwise
// only polls the solver.  Otherwise the phases run per scenario.  Contexts in
// the FP64 form are solved one after the other.
struct PhaseHost {
  int S = 0;
  Dev* d_devs = nullptr;          // scenario tables (swapped in place on the device)
  int* d_pre = nullptr;           // PH_NKIND x (S+1) block prefixes
  double* d_wave = nullptr;       // 3S paddle values of the step
  DevScalars* d_sc = nullptr;     // gathered scalars
  Dev* h_devs = nullptr;          // pinned host staging of the same
  int* h_pre = nullptr;
  double* h_wave = nullptr;
  DevScalars* h_sc = nullptr;
  int64_t launches = 0;           // launches of the last batched step
  // HIP graphs of the fixed-shape phase sequences (before and after the
  // solve), replayed while the launch shapes stay the same
  bool use_graph = false;
  hipGraphExec_t gpre = nullptr, gpost = nullptr;
  int gkey[PH_NKIND] = {-1, -1, -1, -1, -1, -1, -1, -1};
  int64_t gcount[2] = {0, 0};     // kernels in each graph
  int tshape[TK_N] = {TS_P256, TS_P256, TS_P256, TS_P256, TS_P256,
                     TS_P256, TS_P256, TS_P256, TS_P256};  // tuned launch shapes
};

struct isph_batch {
  std::vector<isph_ctx*> ctx;
  std::vector<hipStream_t> own;  // each context's stream, restored on destroy
  hipStream_t stream = nullptr;
  BatchHost* pb = nullptr;       // the precision-form contexts as one solve batch
  std::vector<isph_ctx*> pctx;
  std::vector<isph_solve_stats> last;
  PhaseHost* ph = nullptr;       // batched particle phases (null: per-scenario phases)
};

static void ph_drop_graphs(PhaseHost* h) {
  if (h->gpre) (void)hipGraphExecDestroy(h->gpre);
  if (h->gpost) (void)hipGraphExecDestroy(h->gpost);
  h->gpre = h->gpost = nullptr;
  for (int& k : h->gkey) k = -1;
}

static void ph_free(PhaseHost* h) {
  if (!h) return;
  ph_drop_graphs(h);
  (void)hipFree(h->d_devs);
  (void)hipFree(h->d_pre);
  (void)hipFree(h->d_wave);
  (void)hipFree(h->d_sc);
  (void)hipHostFree(h->h_devs);
  (void)hipHostFree(h->h_pre);
  (void)hipHostFree(h->h_wave);
  (void)hipHostFree(h->h_sc);
  delete h;
}

static int ph_make(int S, PhaseHost** out) {
  PhaseHost* h = new PhaseHost();
  h->S = S;
  const size_t npre = (size_t)PH_NKIND * (S + 1);
  hipError_t e = hipMalloc((void**)&h->d_devs, sizeof(Dev) * S);
  if (e == hipSuccess) e = hipMalloc((void**)&h->d_pre, sizeof(int) * npre);
  if (e == hipSuccess) e = hipMalloc((void**)&h->d_wave, sizeof(double) * 3 * S);
  if (e == hipSuccess) e = hipMalloc((void**)&h->d_sc, sizeof(DevScalars) * S);
  if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_devs, sizeof(Dev) * S);
  if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_pre, sizeof(int) * npre);
  if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_wave, sizeof(double) * 3 * S);
  if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_sc, sizeof(DevScalars) * S);
  if (e != hipSuccess) {
    ph_free(h);
    return ISPH_EHIP;
  }
  *out = h;
  return 0;
}

#define PLAUNCH(kern, grid, ...)                                                       \
  do {                                                                                 \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, c0->stream, __VA_ARGS__);       \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess) {                                                            \
      set_err(c0, "launch of %s failed: %s", #kern, hipGetErrorString(e_));           \
      return ISPH_EHIP;                                                                \
    }                                                                                  \
    h->launches++;                                                                     \
  } while (0)

// the scenario tables as the host holds them (every kernel of the batched
// phases reaches a scenario's arrays and scalars through them, so they are
// written before any of them runs: at batch creation and at every step)
static int ph_upload_devs(isph_batch* b) {
  PhaseHost* h = b->ph;
  isph_ctx* c0 = b->ctx[0];
  for (int s = 0; s < h->S; s++) h->h_devs[s] = b->ctx[s]->D;
  HIPCHK(c0, hipMemcpyAsync(h->d_devs, h->h_devs, sizeof(Dev) * h->S, hipMemcpyHostToDevice,
                            c0->stream));
  return 0;
}

// launch of a tuned particle phase with the batch's block size for it
#define TPLAUNCH(tk, kern, ...)                                                        \
  do {                                                                                 \
    const int b_ = ts_block[h->tshape[tk]];                                            \
    const int kind_ = b_ == 64 ? PH_P64 : (b_ == 128 ? PH_P128 : PH_PART);             \
    hipLaunchKernelGGL(kern, dim3(h->h_pre[kind_ * (h->S + 1) + h->S]), dim3(b_), 0,   \
                       c0->stream, __VA_ARGS__);                                       \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess) {                                                            \
      set_err(c0, "launch of %s failed: %s", #kern, hipGetErrorString(e_));           \
      return ISPH_EHIP;                                                                \
    }                                                                                  \
    h->launches++;                                                                     \
  } while (0)

// the batched neighbour search in the batch's tuned launch shape
static int ph_neighbour(isph_batch* b, const PhaseBatch& B) {
  PhaseHost* h = b->ph;
  isph_ctx* c0 = b->ctx[0];
  const int sh = h->tshape[TK_NEIGHBOUR];
  if (sh == TS_TILE4 || sh == TS_TILE8) {
    const int kind = sh == TS_TILE4 ? PH_T4 : PH_T8;
    const dim3 grid(std::max(1, h->h_pre[kind * (h->S + 1) + h->S]));
    if (sh == TS_TILE4)
      hipLaunchKernelGGL(kb_neighbour_tile4, grid, dim3(128), 0, c0->stream, B);
    else
      hipLaunchKernelGGL(kb_neighbour_tile8, grid, dim3(256), 0, c0->stream, B);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_err(c0, "launch of the tiled neighbour search failed: %s", hipGetErrorString(e));
      return ISPH_EHIP;
    }
    h->launches++;
    return 0;
  }
  TPLAUNCH(TK_NEIGHBOUR, kb_neighbour, B);
  return 0;
}

// one read of every scenario's scalars (the device tables are current)
static int ph_gather(isph_batch* b) {
  PhaseHost* h = b->ph;
  isph_ctx* c0 = b->ctx[0];
  const PhaseBatch B{h->S, h->d_devs, h->d_pre, h->d_wave};
  PLAUNCH(kb_gather_phase, (h->S + 63) / 64, B, h->d_sc);
  HIPCHK(c0, hipMemcpyAsync(h->h_sc, h->d_sc, sizeof(DevScalars) * h->S, hipMemcpyDeviceToHost,
                            c0->stream));
  HIPCHK(c0, hipStreamSynchronize(c0->stream));
  return 0;
}

// the step's tables: each scenario's Dev with its row blocks for the live
// count just read, the block prefixes of its index spaces, its paddle values
static int ph_upload(isph_batch* b) {
  PhaseHost* h = b->ph;
  isph_ctx* c0 = b->ctx[0];
  const int S = h->S;
  for (int k = 0; k < PH_NKIND; k++) h->h_pre[k * (S + 1)] = 0;
  for (int s = 0; s < S; s++) {
    const isph_ctx* c = b->ctx[s];
    h->h_devs[s] = c->D;
    const int nb = c->ncell * NCLS + 1;
    const int blocks[PH_NKIND] = {c->D.nblk, grid_of(c->ncell + 1), grid_of(nb),
                                  (nb + SCAN_TILE - 1) / SCAN_TILE,
                                  std::max(1, (c->n_hi + 63) / 64),
                                  std::max(1, (c->n_hi + 127) / 128),
                                  ntile_count(4, c->D.P.cellx, c->D.P.celly),
                                  ntile_count(8, c->D.P.cellx, c->D.P.celly)};
    for (int k = 0; k < PH_NKIND; k++)
      h->h_pre[k * (S + 1) + s + 1] = h->h_pre[k * (S + 1) + s] + blocks[k];
  }
  HIPCHK(c0, hipMemcpyAsync(h->d_devs, h->h_devs, sizeof(Dev) * S, hipMemcpyHostToDevice,
                            c0->stream));
  HIPCHK(c0, hipMemcpyAsync(h->d_pre, h->h_pre, sizeof(int) * PH_NKIND * (S + 1),
                            hipMemcpyHostToDevice, c0->stream));
  HIPCHK(c0, hipMemcpyAsync(h->d_wave, h->h_wave, sizeof(double) * 3 * S,
                            hipMemcpyHostToDevice, c0->stream));
  return 0;
}

// rebin of every scenario (the counting sort of rebin(), batched)
static int ph_rebin(isph_batch* b, const PhaseBatch& B, int alt) {
  PhaseHost* h = b->ph;
  isph_ctx* c0 = b->ctx[0];
  const int S = h->S;
  auto tot = [&](int kind) { return h->h_pre[kind * (S + 1) + S]; };
  PLAUNCH(kb_fill_bins, tot(PH_BIN), B);
  PLAUNCH(kb_rebin_classify, tot(PH_PART), B, alt);
  PLAUNCH(kb_rebin_rank, tot(PH_PART), B);
  PLAUNCH(kb_scan_tiles, tot(PH_TILE), B);
  PLAUNCH(kb_scan_sums, S, B);
  PLAUNCH(kb_scan_add, tot(PH_BIN), B);
  PLAUNCH(kb_rebin_scatter, tot(PH_PART), B);
  PLAUNCH(kb_swap, (S + 63) / 64, B);
  PLAUNCH(kb_rebin_finish, tot(PH_CELL), B);
  return 0;
}

// the host's half of a batched rebin: the ping-pong swap the device did
static void ph_rebin_host(isph_batch* b) {
  for (isph_ctx* c : b->ctx) {
    swap_state(c->D);
    c->assembled = false;
    c->have_nbr = false;
  }
}

// the phases from projection to the PPE assembly (after the scalar sub-steps)
static int ph_pre_launch(isph_batch* b, const PhaseBatch& B) {
  PhaseHost* h = b->ph;
  isph_ctx* c0 = b->ctx[0];
  const int S = h->S;
  auto tot = [&](int kind) { return h->h_pre[kind * (S + 1) + S]; };
  PLAUNCH(kb_projection, tot(PH_PART), B);
  PLAUNCH(kb_paddle, tot(PH_PART), B);
  int rc = ph_rebin(b, B, 0);
  if (rc == 0) rc = ph_neighbour(b, B);
  if (rc) return rc;
  TPLAUNCH(TK_EFFPOR, kb_effpor, B);
  TPLAUNCH(TK_EDDYVIS, kb_eddyvis, B);
  TPLAUNCH(TK_VARTS, kb_int_vel_varts, B);
  TPLAUNCH(TK_INTVEL, kb_int_vel, B);
  PLAUNCH(kb_zero_nnz, (S + 63) / 64, B);
  TPLAUNCH(TK_ASSEMBLE, kb_ppe_assemble, B);
  return 0;
}

// the phases from comp_vel to the closing eddyvis
static int ph_post_launch(isph_batch* b, const PhaseBatch& B) {
  PhaseHost* h = b->ph;
  isph_ctx* c0 = b->ctx[0];
  const int S = h->S;
  auto tot = [&](int kind) { return h->h_pre[kind * (S + 1) + S]; };
  TPLAUNCH(TK_COMPVEL, kb_comp_vel, B);
  PLAUNCH(kb_cell_max, tot(PH_CELL), B, 0);
  PLAUNCH(kb_comp_pos, tot(PH_PART), B);
  int rc = ph_rebin(b, B, 1);
  if (rc == 0) rc = ph_neighbour(b, B);
  if (rc) return rc;
  PLAUNCH(kb_freesurf, tot(PH_PART), B);
  TPLAUNCH(TK_EFFPOR, kb_effpor, B);
  TPLAUNCH(TK_SHIFT, kb_opt2_shift, B);
  PLAUNCH(kb_shift_apply, tot(PH_PART), B);
  PLAUNCH(kb_densityupdate, tot(PH_PART), B);
  PLAUNCH(kb_timestep, S, B);
  TPLAUNCH(TK_EDDYVIS, kb_eddyvis, B);
  return 0;
}

// Run a phase sequence directly, or through its graph: captured on first
// use and again whenever a launch shape (the block totals) changes, then
// replayed.  The kernels' arguments are the batch's device tables, which
// keep their addresses; what changes between steps lives in them.
static int ph_run(isph_batch* b, const PhaseBatch& B, bool pre) {
  PhaseHost* h = b->ph;
  isph_ctx* c0 = b->ctx[0];
  const int S = h->S;
  if (!h->use_graph) return pre ? ph_pre_launch(b, B) : ph_post_launch(b, B);
  bool same = true;
  for (int k = 0; k < PH_NKIND; k++) same = same && h->gkey[k] == h->h_pre[k * (S + 1) + S];
  if (!same) ph_drop_graphs(h);
  hipGraphExec_t& ex = pre ? h->gpre : h->gpost;
  if (!ex) {
    hipGraph_t g = nullptr;
    const int64_t l0 = h->launches;
    HIPCHK(c0, hipStreamBeginCapture(c0->stream, hipStreamCaptureModeRelaxed));
    int rc = pre ? ph_pre_launch(b, B) : ph_post_launch(b, B);
    h->gcount[pre ? 0 : 1] = h->launches - l0;
    hipError_t e = hipStreamEndCapture(c0->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      set_err(c0, "graph capture: %s", hipGetErrorString(e));
      return ISPH_EHIP;
    }
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      ex = nullptr;
      set_err(c0, "graph instantiation: %s", hipGetErrorString(e));
      return ISPH_EHIP;
    }
    for (int k = 0; k < PH_NKIND; k++) h->gkey[k] = h->h_pre[k * (S + 1) + S];
  } else {
    // a replay: the same launches as the capture
    h->launches += h->gcount[pre ? 0 : 1];
  }
  HIPCHK(c0, hipGraphLaunch(ex, c0->stream));
  return 0;
}

// the scalars gathered at the start of a batched step: the error bits the
// previous step's phases left, and the scalar sub-step counts
static int ph_check(isph_batch* b) {
  PhaseHost* h = b->ph;
  for (int s = 0; s < h->S; s++) {
    isph_ctx* c = b->ctx[s];
    const DevScalars& d = h->h_sc[s];
    if (d.err) {
      int rc = check_err_flags(c);
      if (rc) return rc;
    }
    if (d.solsteps < 1 || d.solsteps > ISPH_MAX_SOLSTEPS) {
      set_err(c, "scalar sub-step count %d outside [1, %d]", d.solsteps, ISPH_MAX_SOLSTEPS);
      return ISPH_ENONFINITE;
    }
    set_nhi(c, d.n);
  }
  return 0;
}

// One production step of every scenario, the particle phases batched: the
// order of step_pre, the solve and step_post (main.f90:199-261).
static int ph_step(isph_batch* b, std::vector<isph_solve_stats>& pst, int* worst) {
  PhaseHost* h = b->ph;
  isph_ctx* c0 = b->ctx[0];
  const int S = h->S;
  h->launches = 0;
  int rc = ph_upload_devs(b);
  if (rc) return rc;
  rc = ph_gather(b);
  if (rc) return rc;
  rc = ph_check(b);
  if (rc) return rc;
  int qmax = 1;
  for (int s = 0; s < S; s++) {
    isph_ctx* c = b->ctx[s];
    const DevScalars& d = h->h_sc[s];
    qmax = std::max(qmax, d.solsteps);
    double X = 0.0, U = 0.0;
    // the device's t + dt, the same double sum
    if (c->wave &&
        isph_wavemaker_motion(c->wv_stroke, c->wv_period, c->wv_ramp, c->wv_t0, d.t + d.dt,
                              &X, &U))
      return ISPH_EARG;
    h->h_wave[3 * s] = X;
    h->h_wave[3 * s + 1] = U;
    h->h_wave[3 * s + 2] = c->wave ? 1.0 : 0.0;
  }
  rc = ph_upload(b);
  if (rc) return rc;
  const PhaseBatch B{S, h->d_devs, h->d_pre, h->d_wave};
  auto tot = [&](int kind) { return h->h_pre[kind * (S + 1) + S]; };
  const int sg = (S + 63) / 64;
  PLAUNCH(kb_advance_t, sg, B);
  for (int q = 0; q < qmax; q++) {
    TPLAUNCH(TK_SCALART, kb_scalart, B, q);
    PLAUNCH(kb_scalarupdate, tot(PH_PART), B, q);
  }
  rc = ph_run(b, B, true);
  if (rc) return rc;
  ph_rebin_host(b);
  for (isph_ctx* c : b->ctx) {
    c->have_nbr = true;
    c->assembled = true;
  }
  // the solves: the precision-form scenarios together, FP64 ones in turn
  if (b->pb) {
    rc = bh_solve(b->pb, pst.data());
    if (rc) return rc;
  }
  size_t q = 0;
  for (size_t s = 0; s < b->ctx.size(); s++) {
    isph_ctx* c = b->ctx[s];
    isph_solve_stats& st = b->last[s];
    int src = 0;
    if (c->P.precision != ISPH_PPE_FP64) {
      st = pst[q++];
      c->last = st;
      if (!std::isfinite(st.rel_residual) || !std::isfinite(st.bnorm)) {
        set_err(c, "non-finite PPE residual");
        src = ISPH_ENONFINITE;
      } else if (!st.converged) {
        set_err(c, "PPE solve did not converge: %d iterations, rel residual %g",
                st.iterations, st.rel_residual);
        src = ISPH_ENOCONV;
      }
    } else {
      src = check_err_flags(c);
      if (src == 0) src = ppe_solve_impl(c, &st);
    }
    if (solve_fatal(src)) return src;
    if (src) *worst = src;
  }
  rc = ph_run(b, B, false);
  if (rc) return rc;
  ph_rebin_host(b);
  for (isph_ctx* c : b->ctx) c->have_nbr = true;
  return 0;
}

// a scenario the batched phases cover: ISPH (not weakly compressible), no
// insertion clock on a buffer zone
static bool ph_coverable(const isph_ctx* c) {
  if (c->wcsph) return false;
  for (const HostZone& z : c->zones)
    if (z.set && z.interval > 0.0) return false;
  return true;
}

extern "C" int isph_batch_create(int32_t n_scen, isph_ctx* const* ctxs, isph_batch** out) {
  if (out) *out = nullptr;
  if (n_scen <= 0 || !ctxs || !out) return ISPH_EARG;
  for (int s = 0; s < n_scen; s++) {
    if (!ctxs[s]) return ISPH_EARG;
    for (int q = 0; q < s; q++)
      if (ctxs[q] == ctxs[s]) return ISPH_EARG;
    if (ctxs[s]->device != ctxs[0]->device) {
      set_err(ctxs[s], "batched contexts must share one device");
      return ISPH_EARG;
    }
    if (ctxs[s]->in_batch) {
      set_err(ctxs[s], "context already belongs to a batch");
      return ISPH_ESTATE;
    }
  }
  isph_batch* b = new isph_batch();
  b->ctx.assign(ctxs, ctxs + n_scen);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return ISPH_EHIP;
  }
  bool phased = true;
  for (isph_ctx* c : b->ctx) {
    (void)hipStreamSynchronize(c->stream);
    b->own.push_back(c->stream);
    c->stream = b->stream;
    c->in_batch = true;
    if (c->P.precision != ISPH_PPE_FP64 && !c->wcsph) b->pctx.push_back(c);
    phased = phased && ph_coverable(c);
  }
  if (!b->pctx.empty()) {
    int rc = bh_make(b->pctx.data(), (int)b->pctx.size(), &b->pb);
    if (rc) {
      isph_batch_destroy(b);
      return rc;
    }
  }
  if (phased) {
    int rc = ph_make(n_scen, &b->ph);
    if (rc == 0) rc = ph_upload_devs(b);
    if (rc == 0) rc = ph_gather(b);  // the live counts before the first step
    if (rc) {
      set_err(b->ctx[0], "batch phase tables: allocation or upload failed");
      isph_batch_destroy(b);
      return rc;
    }
  }
  b->last.resize(n_scen);
  *out = b;
  return 0;
}

extern "C" int isph_batch_step(isph_batch* b, int32_t nsteps) {
  if (!b || nsteps < 0) return ISPH_EARG;
  int worst = 0;
  std::vector<isph_solve_stats> pst(b->pctx.size());
  if (b->ph) {
    for (int k = 0; k < nsteps; k++) {
      int rc = ph_step(b, pst, &worst);
      if (rc) return rc;
    }
    // the last step's error bits
    int rc = ph_upload_devs(b);
    if (rc == 0) rc = ph_gather(b);
    if (rc) return rc;
    for (int s = 0; s < b->ph->S; s++)
      if (b->ph->h_sc[s].err && (rc = check_err_flags(b->ctx[s]))) return rc;
    return worst;
  }
  for (int k = 0; k < nsteps; k++) {
    for (isph_ctx* c : b->ctx) {
      int rc = step_pre(c);
      if (rc) return rc;
    }
    if (b->pb) {
      int rc = bh_solve(b->pb, pst.data());
      if (rc) return rc;
    }
    size_t q = 0;
    for (size_t s = 0; s < b->ctx.size(); s++) {
      isph_ctx* c = b->ctx[s];
      isph_solve_stats& st = b->last[s];
      int rc = 0;
      if (c->wcsph) {
        st = isph_solve_stats{};  // no pressure solve in the weakly compressible form
      } else if (c->P.precision != ISPH_PPE_FP64) {
        st = pst[q++];
        c->last = st;
        if (!std::isfinite(st.rel_residual) || !std::isfinite(st.bnorm)) {
          set_err(c, "non-finite PPE residual");
          rc = ISPH_ENONFINITE;
        } else if (!st.converged) {
          set_err(c, "PPE solve did not converge: %d iterations, rel residual %g",
                  st.iterations, st.rel_residual);
          rc = ISPH_ENOCONV;
        }
      } else {
        rc = ppe_solve_impl(c, &st);
      }
      if (solve_fatal(rc)) return rc;
      if (rc) worst = rc;
    }
    for (isph_ctx* c : b->ctx) {
      int rc = step_post(c, true);
      if (rc) return rc;
    }
  }
  for (isph_ctx* c : b->ctx) {
    int rc = sync(c);
    if (rc) return rc;
  }
  return worst;
}

extern "C" int isph_batch_set_graph(isph_batch* b, int32_t on) {
  if (!b) return ISPH_EARG;
  if (!b->ph) {
    if (on) set_err(b->ctx[0], "the batch's phases run per scenario; no graph to replay");
    return on ? ISPH_ESTATE : 0;
  }
  HIPCHK(b->ctx[0], hipStreamSynchronize(b->stream));
  ph_drop_graphs(b->ph);
  b->ph->use_graph = on != 0;
  return 0;
}

extern "C" int isph_batch_set_shape(isph_batch* b, int32_t kernel, int32_t shape) {
  if (!b || kernel < 0 || kernel >= TK_N || shape < 0 || shape >= ts_count(kernel))
    return ISPH_EARG;
  if (!b->ph) {
    set_err(b->ctx[0], "the batch's phases run per scenario; set its contexts' shapes");
    return ISPH_ESTATE;
  }
  HIPCHK(b->ctx[0], hipStreamSynchronize(b->stream));
  b->ph->tshape[kernel] = shape;
  ph_drop_graphs(b->ph);  // the launch shapes changed
  return 0;
}

// launches of the last batched step (0 when the phases run per scenario)
extern "C" int64_t isph_batch_launches(isph_batch* b) {
  return (b && b->ph) ? b->ph->launches : 0;
}

// live particles of scenario s after the last batched step
extern "C" int isph_batch_count(isph_batch* b, int32_t s, int64_t* n) {
  if (!b || !n || s < 0 || s >= (int32_t)b->ctx.size()) return ISPH_EARG;
  if (b->ph) {
    *n = b->ph->h_sc[s].n;
    return 0;
  }
  DevScalars d;
  int rc = read_scalars(b->ctx[s], &d);
  if (rc) return rc;
  *n = d.n;
  return 0;
}

extern "C" int isph_batch_stats(isph_batch* b, int32_t s, isph_solve_stats* out) {
  if (!b || !out || s < 0 || s >= (int32_t)b->ctx.size()) return ISPH_EARG;
  *out = b->last[s];
  return 0;
}

extern "C" int isph_batch_destroy(isph_batch* b) {
  if (!b) return 0;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (size_t s = 0; s < b->own.size(); s++) {
    b->ctx[s]->stream = b->own[s];
    b->ctx[s]->in_batch = false;
  }
  bh_free(b->pb);
  ph_free(b->ph);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

// ---------------------------------------------------------------------
// allocation-free FP64 solve (isph_ppe_solve_noalloc)
// ---------------------------------------------------------------------
// Scratch of the allocation-free solve, one static device array per GPU:
// six BiCGStab vectors of ISPH_NOALLOC_MAX_ROWS doubles (r, rh, p, t1, s,
// t2), the block partials of two dot products, then the solver scalars and
// the squared right-hand-side norm.  Host code never uses the array's name
// as an address: scratch_base() looks up each device's own copy once, with
// that device current, and every kernel argument is derived from it.
#define NA_ROWS ((size_t)ISPH_NOALLOC_MAX_ROWS)
#define NA_PART (2 * (NA_ROWS / RED_BLOCK + 2))
#define NA_SC_DOUBLES 48
static_assert(sizeof(DevScalars) + sizeof(double) <= NA_SC_DOUBLES * sizeof(double),
              "solver scalars do not fit their scratch slot");
static __device__ __align__(256) double scratch[6 * NA_ROWS + NA_PART + NA_SC_DOUBLES];

#define NA_MAX_DEVICES 64
static std::mutex g_na_mu;
static double* g_na_base[NA_MAX_DEVICES];

static int na_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_noctx_err = buf;
  return code;
}

// this device's copy of the scratch array (looked up on its first use)
static int scratch_base(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return na_fail(ISPH_EHIP, "isph_ppe_solve_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= NA_MAX_DEVICES)
    return na_fail(ISPH_EHIP, "isph_ppe_solve_noalloc: device %d beyond %d", dev, NA_MAX_DEVICES);
  std::lock_guard<std::mutex> g(g_na_mu);
  if (!g_na_base[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return na_fail(ISPH_EHIP, "isph_ppe_solve_noalloc: hipGetSymbolAddress: %s",
                     hipGetErrorString(e));
    g_na_base[dev] = static_cast<double*>(p);
  }
  *out = g_na_base[dev];
  return 0;
}

// the solver scalars start from zero with the row count set
__global__ void k_na_begin(DevScalars* sc, int n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    DevScalars z{};
    z.n = n;
    *sc = z;
  }
}

// every row length within the capacity and every column inside [0, n)
__global__ void __launch_bounds__(256) k_na_check(Dev D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = D.sc->n;
  if (i >= n) return;
  const int L = D.alen[i];
  bool bad = L < 0 || L > ROW_CAP;
  for (int k = 0; !bad && k < L; k++) {
    const int j = D.acol[(size_t)k * D.cap + i];
    bad = j < 0 || j >= n;
  }
  if (bad) atomicOr(&D.sc->err, 1);
}

__global__ void k_na_set_bnorm(DevScalars* sc, double bnorm) {
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->bnorm = bnorm;
}

__global__ void k_na_zero(double* x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

#define NA_LAUNCH(kern, grid, ...)                                                   \
  do {                                                                               \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, __VA_ARGS__);             \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess)                                                            \
      return na_fail(ISPH_EHIP, "isph_ppe_solve_noalloc: launch of %s failed: %s",   \
                     #kern, hipGetErrorString(e_));                                  \
  } while (0)

static int na_read(hipStream_t st, const DevScalars* d_sc, DevScalars* s, double* bn2) {
  hipError_t e = hipMemcpyAsync(s, d_sc, sizeof *s, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && bn2)
    e = hipMemcpyAsync(bn2, reinterpret_cast<const double*>(d_sc) + NA_SC_DOUBLES - 1,
                       sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess)
    return na_fail(ISPH_EHIP, "isph_ppe_solve_noalloc: read-back: %s", hipGetErrorString(e));
  return 0;
}

extern "C" int isph_ppe_solve_noalloc(int32_t n, int32_t ld, const int32_t* row_len,
                                      const int32_t* col, const double* val, const double* rhs,
                                      double* x, double tol, int32_t max_iter,
                                      int32_t max_outer, void* stream,
                                      isph_solve_stats* stats) {
  g_noctx_err.clear();
  if (n < 0 || (size_t)n > NA_ROWS)
    return na_fail(ISPH_EARG, "isph_ppe_solve_noalloc: %d rows outside [0, %d]", n,
                   ISPH_NOALLOC_MAX_ROWS);
  if (ld < n || (int64_t)ld * ROW_CAP >= (1ll << 31))
    return na_fail(ISPH_EARG, "isph_ppe_solve_noalloc: leading dimension %d below %d rows "
                   "or beyond 2^31/%d", ld, n, ROW_CAP);
  if (!row_len || !col || !val || !rhs || !x)
    return na_fail(ISPH_EARG, "isph_ppe_solve_noalloc: null device pointer");
  if (!(tol > 0.0) || max_iter < 0 || max_outer < 0)
    return na_fail(ISPH_EARG, "isph_ppe_solve_noalloc: tol %g, max_iter %d, max_outer %d",
                   tol, max_iter, max_outer);
  isph_solve_stats out{};
  if (n == 0) {
    out.converged = 1;
    if (stats) *stats = out;
    return 0;
  }
  double* base = nullptr;
  int rc = scratch_base(&base);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Dev D{};
  D.P.tol = tol;
  D.P.max_iter = max_iter;
  D.P.max_outer = max_outer;
  D.cap = ld;
  D.nblk = grid_of(n);
  D.alen = const_cast<int32_t*>(row_len);
  D.acol = const_cast<int32_t*>(col);
  D.aval = const_cast<double*>(val);
  D.b = const_cast<double*>(rhs);
  D.sol = x;
  D.r = base;
  D.rh = base + NA_ROWS;
  D.pp = base + 2 * NA_ROWS;
  D.t1 = base + 3 * NA_ROWS;
  D.sv = base + 4 * NA_ROWS;
  D.t2 = base + 5 * NA_ROWS;
  D.part = base + 6 * NA_ROWS;
  D.sc = reinterpret_cast<DevScalars*>(base + 6 * NA_ROWS + NA_PART);
  double* d_bn2 = base + 6 * NA_ROWS + NA_PART + NA_SC_DOUBLES - 1;
  const int rg = row_grid(n), nb = grid_of(n);
  // clear the scalars the previous call left; check the rows; ||b||^2
  NA_LAUNCH(k_na_begin, 1, D.sc, n);
  NA_LAUNCH(k_na_check, grid_of(n), D);
  NA_LAUNCH(k_dot, rg, D, D.b, D.b);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, D, (int)FIN_DOT, nb, d_bn2);
  DevScalars s;
  double bn2 = 0.0;
  rc = na_read(st, D.sc, &s, &bn2);
  if (rc) return rc;
  if (s.err)
    return na_fail(ISPH_EARG, "isph_ppe_solve_noalloc: a row is longer than %d entries or "
                   "has a column outside [0, %d)", ROW_CAP, n);
  const double bnorm = std::sqrt(bn2);
  out.bnorm = bnorm;
  if (bnorm == 0.0 || !std::isfinite(bnorm)) {
    if (bnorm == 0.0) {
      NA_LAUNCH(k_na_zero, grid_of(n), x, n);
      out.converged = 1;
    }
    out.rel_residual = (bnorm == 0.0) ? 0.0 : bnorm;
  } else {
    // the FP64 iteration of isph_ppe_solve: same launches, same chunked
    // polling, same restarts from the true residual after a breakdown
    NA_LAUNCH(k_na_set_bnorm, 1, D.sc, bnorm);
    NA_LAUNCH(k_bicg_init, rg, D);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, D, (int)FIN_INIT, nb,
                       (double*)nullptr);
    rc = na_read(st, D.sc, &s, nullptr);
    if (rc) return rc;
    int restarts = 0, chunk = 4;
    for (int64_t it = 0; !s.done;) {
      for (int k = 0; k < chunk; k++) {
        NA_LAUNCH(k_bicg_1, rg, D);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, D, (int)FIN_ALPHA, nb,
                           (double*)nullptr);
        NA_LAUNCH(k_bicg_2, rg, D);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, D, (int)FIN_OMEGA, nb,
                           (double*)nullptr);
        NA_LAUNCH(k_bicg_3, rg, D);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, D, (int)FIN_BETA, nb,
                           (double*)nullptr);
        NA_LAUNCH(k_bicg_4, rg, D);
      }
      it += chunk;
      rc = na_read(st, D.sc, &s, nullptr);
      if (rc) return rc;
      if (s.done) {
        if (s.brk && !s.conv && s.iters < max_iter && restarts < max_outer) {
          restarts++;
          NA_LAUNCH(k_bicg_init, rg, D);
          hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, D, (int)FIN_RESTART, nb,
                             (double*)nullptr);
          rc = na_read(st, D.sc, &s, nullptr);
          if (rc) return rc;
          chunk = 4;
          continue;
        }
        break;
      }
      if (chunk < 32) chunk *= 2;
      if (it > (int64_t)(max_iter + 64) * (max_outer + 1)) break;
    }
    out.iterations = s.iters;
    out.outer = restarts;
    out.breakdown = s.brk;
    out.converged = s.conv;
    out.rel_residual = s.rel;
  }
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess)
    return na_fail(ISPH_EHIP, "isph_ppe_solve_noalloc: %s", hipGetErrorString(e));
  if (stats) *stats = out;
  if (!std::isfinite(out.rel_residual) || !std::isfinite(out.bnorm))
    return na_fail(ISPH_ENONFINITE, "isph_ppe_solve_noalloc: non-finite residual");
  if (!out.converged)
    return na_fail(ISPH_ENOCONV, "isph_ppe_solve_noalloc: %d iterations, rel residual %g",
                   out.iterations, out.rel_residual);
  return 0;
}

// ---------------------------------------------------------------------
// autotuner: launch shapes of the particle kernels, measured on the state
// ---------------------------------------------------------------------
// Mean microseconds per launch of `launch` over `trials` launches after one
// untimed launch, from HIP events on the stream.
template <typename F>
static int tune_time(isph_ctx* c, hipStream_t st, int trials, F launch, double* us) {
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  int rc = launch();
  if (rc == 0) HIPCHK(c, hipEventRecord(e0, st));
  for (int k = 0; rc == 0 && k < trials; k++) rc = launch();
  if (rc == 0) HIPCHK(c, hipEventRecord(e1, st));
  if (rc == 0) HIPCHK(c, hipEventSynchronize(e1));
  float ms = 0.f;
  if (rc == 0) HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *us = 1e3 * ms / trials;
  return rc;
}

extern "C" int isph_set_shape(isph_ctx* c, int32_t kernel, int32_t shape) {
  if (!c || kernel < 0 || kernel >= TK_N || shape < 0 || shape >= ts_count(kernel))
    return ISPH_EARG;
  c->tshape[kernel] = shape;
  return 0;
}

extern "C" int isph_autotune(isph_ctx* c, int32_t trials, int32_t* shape, double* us) {
  if (!c || trials < 1) return ISPH_EARG;
  if (!c->have_nbr || c->wcsph) {
    set_err(c, "autotune needs an ISPH state at a step boundary (after init_state or a step)");
    return ISPH_ESTATE;
  }
  DevScalars sc;
  int rc = read_scalars(c, &sc);
  if (rc) return rc;
  set_nhi(c, sc.n);
  Dev& D = c->D;
  const size_t nb = 8 * (size_t)sc.n;
  // comp_vel rewrites u and v: keep them in the other half of the ping-pong
  // state (free until the next rebin writes it) and put them back after
  HIPCHK(c, hipMemcpyAsync(D.u2, D.u, nb, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(D.v2, D.v, nb, hipMemcpyDeviceToDevice, c->stream));
  for (int tk = 0; tk < TK_N; tk++) {
    double best = 1e300;
    int sbest = TS_P256;
    for (int q = 0; q < TS_N; q++) {
      if (q >= ts_count(tk)) {
        if (us) us[TS_N * tk + q] = -1.0;
        continue;
      }
      c->tshape[tk] = q;
      auto launch = [&]() -> int {
        switch (tk) {
          case TK_NEIGHBOUR: return launch_neighbour(c);
          case TK_EFFPOR: TLAUNCH(tk, k_effpor, c->n_hi, c->D); break;
          case TK_EDDYVIS: TLAUNCH(tk, k_eddyvis, c->n_hi, c->D); break;
          case TK_VARTS: TLAUNCH(tk, k_int_vel_varts, c->n_hi, c->D); break;
          case TK_INTVEL: TLAUNCH(tk, k_int_vel, c->n_hi, c->D); break;
          case TK_ASSEMBLE: TLAUNCH(tk, k_ppe_assemble, c->n_hi, c->D); break;
          case TK_COMPVEL: TLAUNCH(tk, k_comp_vel, c->n_hi, c->D); break;
          case TK_SHIFT: TLAUNCH(tk, k_opt2_shift, c->n_hi, c->D); break;
          default: TLAUNCH(tk, k_scalart, c->n_hi, c->D); break;
        }
        return 0;
      };
      double t = 0.0;
      rc = tune_time(c, c->stream, trials, launch, &t);
      if (rc) return rc;
      if (us) us[TS_N * tk + q] = t;
      if (t < best) {
        best = t;
        sbest = q;
      }
    }
    c->tshape[tk] = sbest;
    if (shape) shape[tk] = sbest;
  }
  HIPCHK(c, hipMemcpyAsync(D.u, D.u2, nb, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(D.v, D.v2, nb, hipMemcpyDeviceToDevice, c->stream));
  return check_err_flags(c);
}

extern "C" int isph_batch_autotune(isph_batch* b, int32_t trials, int32_t* shape, double* us) {
  if (!b || trials < 1) return ISPH_EARG;
  isph_ctx* c0 = b->ctx[0];
  if (!b->ph) {
    set_err(c0, "the batch's phases run per scenario; tune its contexts instead");
    return ISPH_ESTATE;
  }
  for (isph_ctx* c : b->ctx)
    if (!c->have_nbr) {
      set_err(c0, "autotune needs every scenario at a step boundary");
      return ISPH_ESTATE;
    }
  PhaseHost* h = b->ph;
  int rc = ph_upload_devs(b);
  if (rc == 0) rc = ph_gather(b);
  if (rc == 0) rc = ph_check(b);
  for (int s = 0; rc == 0 && s < h->S; s++) h->h_wave[3 * s + 2] = 0.0;
  if (rc == 0) rc = ph_upload(b);
  if (rc) return rc;
  for (isph_ctx* c : b->ctx) {
    const size_t nb = 8 * (size_t)c->n_hi;
    HIPCHK(c0, hipMemcpyAsync(c->D.u2, c->D.u, nb, hipMemcpyDeviceToDevice, c0->stream));
    HIPCHK(c0, hipMemcpyAsync(c->D.v2, c->D.v, nb, hipMemcpyDeviceToDevice, c0->stream));
  }
  const PhaseBatch B{h->S, h->d_devs, h->d_pre, h->d_wave};
  for (int tk = 0; tk < TK_N; tk++) {
    double best = 1e300;
    int sbest = TS_P256;
    for (int q = 0; q < TS_N; q++) {
      if (q >= ts_count(tk)) {
        if (us) us[TS_N * tk + q] = -1.0;
        continue;
      }
      h->tshape[tk] = q;
      auto launch = [&]() -> int {
        switch (tk) {
          case TK_NEIGHBOUR: return ph_neighbour(b, B);
          case TK_EFFPOR: TPLAUNCH(tk, kb_effpor, B); break;
          case TK_EDDYVIS: TPLAUNCH(tk, kb_eddyvis, B); break;
          case TK_VARTS: TPLAUNCH(tk, kb_int_vel_varts, B); break;
          case TK_INTVEL: TPLAUNCH(tk, kb_int_vel, B); break;
          case TK_ASSEMBLE: TPLAUNCH(tk, kb_ppe_assemble, B); break;
          case TK_COMPVEL: TPLAUNCH(tk, kb_comp_vel, B); break;
          case TK_SHIFT: TPLAUNCH(tk, kb_opt2_shift, B); break;
          default: TPLAUNCH(tk, kb_scalart, B, 0); break;
        }
        return 0;
      };
      double t = 0.0;
      rc = tune_time(c0, c0->stream, trials, launch, &t);
      if (rc) return rc;
      if (us) us[TS_N * tk + q] = t;
      if (t < best) {
        best = t;
        sbest = q;
      }
    }
    h->tshape[tk] = sbest;
    if (shape) shape[tk] = sbest;
  }
  for (isph_ctx* c : b->ctx) {
    const size_t nb = 8 * (size_t)c->n_hi;
    HIPCHK(c0, hipMemcpyAsync(c->D.u, c->D.u2, nb, hipMemcpyDeviceToDevice, c0->stream));
    HIPCHK(c0, hipMemcpyAsync(c->D.v, c->D.v2, nb, hipMemcpyDeviceToDevice, c0->stream));
  }
  ph_drop_graphs(h);  // the launch shapes changed
  rc = ph_gather(b);
  if (rc) return rc;
  for (int s = 0; s < h->S; s++)
    if (h->h_sc[s].err && (rc = check_err_flags(b->ctx[s]))) return rc;
  return 0;
}
