// pepsmg_core.hip -- hierarchy, V-cycle orchestration and the C-ABI of
// libpepsmg.so.  See include/pepsmg.h for the interface contract.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pepsmg.h"
#include "pepsmg_internal.h"

namespace pepsmg {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(PEPSMG_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

long pitch_for(int nx) { return ((long)nx + 15) / 16 * 16; }

// ---------------------------------------------------------------------------
// Device buffers
template <typename T>
static T* dalloc(size_t n, int* rc) {
  void* p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) {
    *rc = fail(PEPSMG_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
    return nullptr;
  }
  return (T*)p;
}

template <typename T>
static T* dupload(const std::vector<T>& v, int* rc) {
  T* p = dalloc<T>(v.size(), rc);
  if (!p) return nullptr;
  hipError_t e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    *rc = hip_fail(e, "hipMemcpy H2D");
    return nullptr;
  }
  return p;
}

template <typename S, typename T>
static std::vector<T> conv(const std::vector<S>& v) {
  std::vector<T> o(v.size());
  for (size_t q = 0; q < v.size(); q++) o[q] = (T)v[q];
  return o;
}

// Coefficient arrays of one axis (src/matrix.f90:68-80): cw = 1/(dm*dg(i)),
// ce = 1/(dm*dg(i+1)), index 1..n, zero at 0 and n+1.
static void axis_coeffs(const std::vector<double>& dm, const std::vector<double>& dg, std::vector<double>& cw,
                        std::vector<double>& ce) {
  size_t n = dm.size() - 2;
  cw.assign(n + 2, 0.0);
  ce.assign(n + 2, 0.0);
  for (size_t i = 1; i <= n; i++) {
    cw[i] = 1.0 / (dm[i] * dg[i]);
    ce[i] = 1.0 / (dm[i] * dg[i + 1]);
  }
}

// Coarse metrics of one axis (multigrid_subdomain_CGS_make_grid,
// src/multigrid.f90:445-493).  xg is not used by the solver and not kept.
static void make_grid(std::vector<double>& dm, std::vector<double>& dg, int n, const std::vector<double>& dmf,
                      const std::vector<double>& dgf, int lv_cur, int lv_coarsest) {
  dm.assign(n + 2, 0.0);
  dg.assign(n + 2, 0.0);
  if (lv_cur <= lv_coarsest) {
    for (int i = 1; i <= n; i++) dm[i] = dmf[2 * i - 1] + dmf[2 * i];
    dm[0] = dmf[0];
    dm[n + 1] = dmf[2 * n + 1];
    for (int i = 1; i <= n; i++) dg[i] = 0.5 * (dm[i] + dm[i - 1]);
    dg[n + 1] = 0.5 * (dm[n + 1] + dm[n]);
  } else {
    dm = dmf;
    dg = dgf;
  }
}

int HostLevel::setup_coeffs() {
  for (int a = 0; a < 3; a++) axis_coeffs(dm[a], dg[a], cw[a], ce[a]);
  return PEPSMG_OK;
}

// Restriction weights from level f to the next (src/multigrid.f90:1415-1458)
static void restrict_weights(const std::vector<double>& dmf, const std::vector<double>& dmc, bool coarsen,
                             std::vector<double>& w, int& s) {
  size_t nf = dmf.size() - 2;
  if (!coarsen) {
    w = dmf;
    s = 1;
    return;
  }
  w.assign(nf + 2, 0.0);
  for (size_t i = 1; i <= nf; i++) w[i] = dmc[(i - 1) / 2 + 1] - dmf[i];
  s = 2;
}

// Prolongation weights onto level f (src/multigrid.f90:1594-1649)
static void prolong_weights(const std::vector<double>& dmf, const std::vector<double>& dmc, bool coarsen,
                            std::vector<double>& p, std::vector<double>& n, int& s) {
  size_t nf = dmf.size() - 2;
  if (!coarsen) {
    p = dmc;
    n = dmc;
    s = 1;
    return;
  }
  size_t nc = dmc.size() - 2;
  p.assign(nf + 2, 0.0);
  n.assign(nf + 2, 0.0);
  for (size_t i = 1; i <= nc; i++) {
    size_t f = 2 * i;
    p[f - 1] = 1.0 * dmf[f - 2] + 0.5 * dmf[f - 1];
    p[f] = 0.5 * dmf[f - 1];
    n[f - 1] = 0.5 * dmf[f];
    n[f] = 1.0 * dmf[f + 1] + 0.5 * dmf[f];
  }
  s = 2;
}

// Grid descriptor: sizes, metric arrays, and every width dm(1..n) and centre
// distance dg(1..n+1) positive and finite (the coefficients divide by them,
// src/matrix.f90:68-80).  The reference's stretched-axis formula sets cells
// 1..n/2 and their mirrors only (src/geometry.f90:190-197), so an odd cell
// count leaves the middle width unset.
int check_grid(const pepsmg_grid* g) {
  if (!g || g->nx < 1 || g->ny < 1 || g->nz < 1 || !g->dxm || !g->dxg || !g->dym || !g->dyg || !g->dzm || !g->dzg)
    return fail(PEPSMG_ERR_ARG, "invalid grid descriptor");
  const int n[3] = {g->nx, g->ny, g->nz};
  const double* dm[3] = {g->dxm, g->dym, g->dzm};
  const double* dg[3] = {g->dxg, g->dyg, g->dzg};
  for (int a = 0; a < 3; a++) {
    for (int i = 1; i <= n[a] + 1; i++) {
      const bool ok_m = i > n[a] || (std::isfinite(dm[a][i]) && dm[a][i] > 0.0);
      const bool ok_g = std::isfinite(dg[a][i]) && dg[a][i] > 0.0;
      if (!ok_m || !ok_g)
        return fail(PEPSMG_ERR_ARG,
                    "grid metric %s(%d) of the %c axis is not positive and finite (a stretched axis needs an even "
                    "cell count)",
                    ok_m ? "dg" : "dm", i, "xyz"[a]);
    }
  }
  return PEPSMG_OK;
}

// ---------------------------------------------------------------------------
// Hierarchy (multigrid_create and the 1-process level rules)
int build_hierarchy(const pepsmg_grid* g, int nlevel, int aggr, int cga_level, std::vector<HostLevel>& lv,
                    int lvc[3]) {
  if (int rc = check_grid(g)) return rc;
  if (nlevel <= 1)
    return fail(PEPSMG_ERR_LEVELS, "[Error] The number of level should be larger than 1. Current number: %d", nlevel);
  if (aggr < 0 || aggr > 2)
    return fail(PEPSMG_ERR_LEVELS, "[Error] Aggregation method should be 0, 1, or 2. %d", aggr);
  if (aggr == 0 && cga_level != 0)
    return fail(PEPSMG_ERR_LEVELS, "[Error] Aggretation level should be 0 for CGS method.");
  if (aggr == 1 && cga_level == 0)
    return fail(PEPSMG_ERR_LEVELS, "[Error] Aggretation level should be larger than 0");
  if (nlevel > PEPSMG_MAX_LEVELS) return fail(PEPSMG_ERR_LEVELS, "at most %d levels supported", PEPSMG_MAX_LEVELS);
  int n[3] = {g->nx, g->ny, g->nz};
  lvc[0] = lvc[1] = lvc[2] = 0;
  lv.assign(nlevel + 1, HostLevel());
  lv[0].n[0] = n[0]; lv[0].n[1] = n[1]; lv[0].n[2] = n[2];
  const double* dms[3] = {g->dxm, g->dym, g->dzm};
  const double* dgs[3] = {g->dxg, g->dyg, g->dzg};
  for (int a = 0; a < 3; a++) {
    lv[0].dm[a].assign(dms[a], dms[a] + n[a] + 2);
    lv[0].dg[a].assign(dgs[a], dgs[a] + n[a] + 2);
  }
  bool stopped = false;
  int made = 0;
  for (int l = 1; l <= nlevel && !stopped; l++) {
    for (int a = 0; a < 3; a++) {
      if (n[a] % 2 == 0) {
        n[a] /= 2;
        lvc[a] = l;
      } else if (aggr == 1 && l <= cga_level) {
        return fail(PEPSMG_ERR_LEVELS, "[Error] %c-grid coarsening impossible at level = %d, <= aggregation level = %d",
                    "XYZ"[a], l, cga_level);
      }
    }
    for (int a = 0; a < 3; a++) lv[l].n[a] = n[a];
    made = l;
    if (aggr != 0 && ((long)n[0] * n[1] * n[2]) % 2 == 1) stopped = true;
  }
  int lvmax = std::max(lvc[0], std::max(lvc[1], lvc[2]));
  if (aggr == 2 ? (nlevel != lvmax) : (nlevel > lvmax) || made < nlevel)
    return fail(PEPSMG_ERR_LEVELS,
                aggr == 2 ? "[MG] Number of levels should be equal to the coarsest level in global domain (%d vs %d)."
                          : "[MG] Number of levels is greater than the coarsest level in global domain (%d > %d).",
                nlevel, lvmax);
  for (int l = 1; l <= nlevel; l++)
    for (int a = 0; a < 3; a++)
      make_grid(lv[l].dm[a], lv[l].dg[a], lv[l].n[a], lv[l - 1].dm[a], lv[l - 1].dg[a], l, lvc[a]);
  for (int l = 0; l <= nlevel; l++) {
    lv[l].setup_coeffs();
    if (l < nlevel)
      for (int a = 0; a < 3; a++) {
        bool co = l < lvc[a];
        restrict_weights(lv[l].dm[a], lv[l + 1].dm[a], co, lv[l].rw[a], lv[l].rs[a]);
        prolong_weights(lv[l].dm[a], lv[l + 1].dm[a], co, lv[l].pp[a], lv[l].pn[a], lv[l].ps[a]);
      }
  }
  return PEPSMG_OK;
}

// ---------------------------------------------------------------------------
template <typename T>
static int upload_level(const HostLevel& h, DevLevel<T>& d, bool with_transfer) {
  int rc = PEPSMG_OK;
  d.nx = h.n[0]; d.ny = h.n[1]; d.nz = h.n[2];
  d.px = pitch_for(d.nx);
  d.pxy = d.px * d.ny;
  d.ncell = (size_t)d.pxy * d.nz;
  for (int a = 0; a < 3; a++) {
    d.cw[a] = dupload(conv<double, T>(h.cw[a]), &rc);
    d.ce[a] = dupload(conv<double, T>(h.ce[a]), &rc);
    if (with_transfer) {
      d.rw[a] = dupload(conv<double, T>(h.rw[a]), &rc);
      d.pp[a] = dupload(conv<double, T>(h.pp[a]), &rc);
      d.pn[a] = dupload(conv<double, T>(h.pn[a]), &rc);
      d.rs[a] = h.rs[a];
      d.ps[a] = h.ps[a];
    }
  }
  if (rc) return rc;
  d.x = dalloc<T>(d.ncell, &rc);
  d.b = dalloc<T>(d.ncell, &rc);
  d.r = dalloc<T>(d.ncell, &rc);
  d.x2 = dalloc<T>(d.ncell, &rc);
  if (rc) return rc;
  hipMemset(d.x, 0, d.ncell * sizeof(T));
  hipMemset(d.b, 0, d.ncell * sizeof(T));
  hipMemset(d.r, 0, d.ncell * sizeof(T));
  hipMemset(d.x2, 0, d.ncell * sizeof(T));
  return PEPSMG_OK;
}

template <typename T>
static void free_level(DevLevel<T>& d) {
  for (int a = 0; a < 3; a++) {
    hipFree(d.cw[a]); hipFree(d.ce[a]); hipFree(d.rw[a]); hipFree(d.pp[a]); hipFree(d.pn[a]);
    d.cw[a] = d.ce[a] = d.rw[a] = d.pp[a] = d.pn[a] = nullptr;
  }
  hipFree(d.x); hipFree(d.b); hipFree(d.r); hipFree(d.x2);
  d.x = d.b = d.r = d.x2 = nullptr;
}

template <typename T>
int Hier<T>::build(const std::vector<HostLevel>& hl) {
  lv.assign(hl.size(), DevLevel<T>());
  for (size_t l = 0; l < hl.size(); l++) {
    int rc = upload_level(hl[l], lv[l], l + 1 < hl.size());
    if (rc) return rc;
  }
  return PEPSMG_OK;
}

template <typename T>
void Hier<T>::release() {
  for (auto& d : lv) free_level(d);
  lv.clear();
}

template struct Hier<double>;
template struct Hier<float>;

// Synthetic right-hand side: SplitMix64 (Steele, Lea and Flood 2014) of
// seed ^ linear index, top 53 bits as a uniform in [0,1).
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_fill_rhs(double* __restrict__ out, int nx, int ny, int nz, uint64_t seed, double sr, double ss,
                           const double* __restrict__ xg, const double* __restrict__ yg,
                           const double* __restrict__ zg) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i > nx + 1 || j > ny + 1) return;
  long s = (long)i + (long)(nx + 2) * ((long)j + (long)(ny + 2) * k);
  if (i == 0 || j == 0 || k == 0 || i == nx + 1 || j == ny + 1 || k == nz + 1) {
    out[s] = 0.0;
    return;
  }
  uint64_t lin = (uint64_t)(i - 1) + (uint64_t)nx * ((uint64_t)(j - 1) + (uint64_t)ny * (uint64_t)(k - 1));
  double u = (double)(splitmix64(seed ^ lin) >> 11) * (1.0 / 9007199254740992.0);
  const double pi = 3.14159265358979323846;
  double v = sr * (2.0 * u - 1.0);
  if (ss != 0.0) v = v + ss * (sin(pi * xg[i]) * sin(pi * yg[j]) * sin(pi * zg[k]));
  out[s] = v;
}

}  // namespace pepsmg

// ---------------------------------------------------------------------------
// C-ABI
using namespace pepsmg;

extern "C" {

const char* pepsmg_last_error(void) { return g_err.c_str(); }

void pepsmg_default_options(pepsmg_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->smoother = 0;
  o->precision = 0;
  o->coarse_switch = 4096;
  o->inner_cycles = 2;
  o->max_refine = 30;
  o->fused_smoother = 1;
  o->tile_y = 0;
  o->verbose = 0;
  o->tile_x = 0;
  o->inner_nu = 2;
  o->inner_omega = 1.35;
  o->inner_nu_post = 0;
}

int pepsmg_multigrid_create(const pepsmg_grid* grid, int32_t nlevel, int32_t ncycle, int32_t aggr_method,
                            int32_t single_aggr_level, const int32_t adaptive_aggr_level[3], pepsmg_ctx** out) {
  (void)adaptive_aggr_level;  // on one process CGPSA sets its aggregation levels to 0 (src/multigrid.f90:1032-1036)
  if (!out) return fail(PEPSMG_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (ncycle < 0) return fail(PEPSMG_ERR_ARG, "ncycle < 0");
  pepsmg_ctx* c = new pepsmg_ctx();
  int rc = build_hierarchy(grid, nlevel, aggr_method, single_aggr_level, c->host, c->lvc);
  if (rc) {
    delete c;
    return rc;
  }
  c->nlevel = nlevel;
  c->ncycle = ncycle;
  pepsmg_default_options(&c->opt);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  {
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 4096)
      c->lds_max = (size_t)lds - 2048;   // room for the kernel's static LDS
  }
  rc = c->h64.build(c->host);
  if (rc == PEPSMG_OK) rc = c->alloc_scratch();
  if (rc) {
    c->release();
    delete c;
    return rc;
  }
  *out = c;
  return PEPSMG_OK;
}

int pepsmg_set_options(pepsmg_ctx* ctx, const pepsmg_options* opt) {
  if (!ctx || !opt) return fail(PEPSMG_ERR_ARG, "NULL argument");
  if (opt->smoother < 0 || opt->smoother > 1) return fail(PEPSMG_ERR_ARG, "smoother must be 0 or 1");
  if (opt->precision < 0 || opt->precision > 1) return fail(PEPSMG_ERR_ARG, "precision must be 0 or 1");
  if (opt->inner_cycles < 1 || opt->max_refine < 1) return fail(PEPSMG_ERR_ARG, "inner_cycles/max_refine must be >= 1");
  if (opt->fused_smoother < 0 || opt->fused_smoother > 2) return fail(PEPSMG_ERR_ARG, "fused_smoother must be 0, 1 or 2");
  if (opt->tile_y != 0 && opt->tile_y != 8 && opt->tile_y != 16 && opt->tile_y != 32)
    return fail(PEPSMG_ERR_ARG, "tile_y must be 0, 8, 16 or 32");
  if (opt->tile_x != 0 && opt->tile_x != 64 && opt->tile_x != 128) return fail(PEPSMG_ERR_ARG, "tile_x must be 0, 64 or 128");
  if (opt->inner_nu < 0 || opt->inner_nu_post < 0) return fail(PEPSMG_ERR_ARG, "inner_nu/inner_nu_post must be >= 0");
  if (!(opt->inner_omega >= 0.0 && opt->inner_omega < 2.0))
    return fail(PEPSMG_ERR_ARG, "inner_omega must be 0 (caller's omega) or in (0, 2)");
  ctx->opt = *opt;
  if (opt->precision == 1 && ctx->h32.lv.empty()) {
    int rc = ctx->h32.build(ctx->host);
    if (rc) return rc;
  }
  return PEPSMG_OK;
}

void pepsmg_multigrid_destroy(pepsmg_ctx* ctx) {
  if (!ctx) return;
  ctx->release();
  delete ctx;
}

int pepsmg_fill_rhs_host(int32_t nx, int32_t ny, int32_t nz, double* out, uint64_t seed, double scale_rand,
                         double scale_sin, const double* xg, const double* yg, const double* zg) {
  if (nx < 1 || ny < 1 || nz < 1 || !out) return fail(PEPSMG_ERR_ARG, "invalid argument");
  if (scale_sin != 0.0 && (!xg || !yg || !zg)) return fail(PEPSMG_ERR_ARG, "cell centres required");
  const double pi = 3.14159265358979323846;
  size_t sx = (size_t)nx + 2, sy = (size_t)ny + 2;
  std::memset(out, 0, sx * sy * ((size_t)nz + 2) * sizeof(double));
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        uint64_t lin = (uint64_t)(i - 1) + (uint64_t)nx * ((uint64_t)(j - 1) + (uint64_t)ny * (uint64_t)(k - 1));
        uint64_t z = seed ^ lin;
        z += 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
        double v = scale_rand * (2.0 * u - 1.0);
        if (scale_sin != 0.0) v = v + scale_sin * (std::sin(pi * xg[i]) * std::sin(pi * yg[j]) * std::sin(pi * zg[k]));
        out[(size_t)i + sx * ((size_t)j + sy * (size_t)k)] = v;
      }
  return PEPSMG_OK;
}

// y = x / 2 + b over n doubles, 16 bytes per lane, output stored
// non-temporally: the access pattern of one smoother launch (two streams in,
// one out) without the stencil, i.e. the bandwidth such a kernel can reach.
__global__ void __launch_bounds__(256) k_stream_triad(double* __restrict__ y, const double* __restrict__ x,
                                                      const double* __restrict__ b, long n2) {
  typedef double v2 __attribute__((ext_vector_type(2)));
  long k = (long)blockIdx.x * 256 + threadIdx.x;
  if (k < n2) {
    v2 a = reinterpret_cast<const v2*>(x)[k], c = reinterpret_cast<const v2*>(b)[k];
    v2 r = a * 0.5 + c;
    __builtin_nontemporal_store(r, reinterpret_cast<v2*>(y) + k);
  }
}

int pepsmg_bench_stream(int64_t n, int32_t reps, double* avg_ms, double* bytes) {
  if (n < 2 || reps < 1) return fail(PEPSMG_ERR_ARG, "invalid argument");
  n &= ~1LL;
  double *x = nullptr, *b = nullptr, *y = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PEPSMG_OK;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == PEPSMG_OK) rc = hip_fail(e, what);
  };
  chk(hipMalloc(&x, n * sizeof(double)), "hipMalloc");
  chk(hipMalloc(&b, n * sizeof(double)), "hipMalloc");
  chk(hipMalloc(&y, n * sizeof(double)), "hipMalloc");
  chk(hipEventCreate(&e0), "hipEventCreate");
  chk(hipEventCreate(&e1), "hipEventCreate");
  if (rc == PEPSMG_OK) {
    chk(hipMemset(x, 0, n * sizeof(double)), "hipMemset");
    chk(hipMemset(b, 0, n * sizeof(double)), "hipMemset");
    long n2 = n / 2;
    dim3 g((unsigned)((n2 + 255) / 256));
    for (int q = 0; q < reps; q++) hipLaunchKernelGGL(k_stream_triad, g, dim3(256), 0, 0, y, x, b, n2);
    chk(hipEventRecord(e0, 0), "hipEventRecord");
    for (int q = 0; q < reps; q++) hipLaunchKernelGGL(k_stream_triad, g, dim3(256), 0, 0, y, x, b, n2);
    chk(hipEventRecord(e1, 0), "hipEventRecord");
    chk(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    chk(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    if (avg_ms) *avg_ms = ms / reps;
    if (bytes) *bytes = 3.0 * n * sizeof(double);
  }
  hipFree(x);
  hipFree(b);
  hipFree(y);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  return rc;
}

int pepsmg_fill_rhs_device(int32_t nx, int32_t ny, int32_t nz, double* d_out, uint64_t seed, double scale_rand,
                           double scale_sin, const double* xg, const double* yg, const double* zg, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1 || !d_out) return fail(PEPSMG_ERR_ARG, "invalid argument");
  if (scale_sin != 0.0 && (!xg || !yg || !zg)) return fail(PEPSMG_ERR_ARG, "cell centres required");
  hipStream_t s = (hipStream_t)stream;
  double* dg = nullptr;
  size_t tot = (size_t)(nx + 2) + (ny + 2) + (nz + 2);
  HIP_TRY(hipMalloc(&dg, tot * sizeof(double)));
  std::vector<double> h(tot, 0.0);
  if (scale_sin != 0.0) {
    std::copy(xg, xg + nx + 2, h.begin());
    std::copy(yg, yg + ny + 2, h.begin() + (nx + 2));
    std::copy(zg, zg + nz + 2, h.begin() + (nx + 2) + (ny + 2));
  }
  HIP_TRY(hipMemcpy(dg, h.data(), tot * sizeof(double), hipMemcpyHostToDevice));
  dim3 g((unsigned)((nx + 2 + 63) / 64), (unsigned)((ny + 2 + 3) / 4), (unsigned)(nz + 2));
  hipLaunchKernelGGL(k_fill_rhs, g, dim3(64, 4, 1), 0, s, d_out, nx, ny, nz, (uint64_t)seed, scale_rand, scale_sin,
                     dg, dg + (nx + 2), dg + (nx + 2) + (ny + 2));
  hipError_t e = hipStreamSynchronize(s);
  hipFree(dg);
  if (e != hipSuccess) return hip_fail(e, "fill_rhs");
  return PEPSMG_OK;
}

}  // extern "C"
