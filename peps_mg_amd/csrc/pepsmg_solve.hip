// pepsmg_solve.hip -- kernel launchers, the V-cycle schedule (FP64 and FP32
// inside FP64 iterative refinement), the CG smoother and the solve /
// standalone-operator entry points of the C-ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "../../include/pepsmg.h"
#include "pepsmg_internal.h"
#include "pepsmg_ops.h"

namespace pepsmg {

// ---------------------------------------------------------------------------
// CG iterator (src/cg_poisson_matrix.f90:157-241) on device arrays: Ap = A p
// with the partial sums of (Ap, p) (one z-march pass), x += alpha p and
// r -= alpha Ap with the partial sums of (r, r) (one elementwise pass), and
// the final reductions; the direction update p = r + beta p of an iteration
// is folded into the next iteration's A p pass (k_cg_ap_fold; after the
// last iteration it is not needed: p is local to the reference routine).
// Scalars stay on the device: sc[it & 1] = (r, r) at the start of iteration
// it (temp1), sc[2] = (Ap, p) (temp2); alpha and beta are formed in FP64 by
// every thread as the reference does (:209, :223).
constexpr int CG_WAVES = 4;

// rows (j, k) of a level dealt to waves; each lane handles pairs (i odd, i+1)
template <typename T, typename F>
__device__ __forceinline__ double cg_rows(const Lvl<T>& L, F&& f) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nrow = L.ny * L.nz, npair = (L.nx + 1) / 2;
  double acc = 0.0;
  for (int row = blockIdx.x * CG_WAVES + w; row < nrow; row += gridDim.x * CG_WAVES) {
    long base = (long)(row / L.ny) * L.pxy + (long)(row % L.ny) * L.px;
    for (int pp = lane; pp < npair; pp += 64) acc += f(base + 2 * pp, 2 * pp + 2 <= L.nx);
  }
  return acc;
}

// Whole pairs (i odd, i+1) of the solver's layout are 16-byte (FP64) or
// 8-byte (FP32) aligned: rows start at multiples of the pitch (a multiple of
// 16 elements) and a pair at an even offset.  The elementwise CG kernels
// move them as one vector per array, so every store instruction writes
// whole cache lines (two interleaved scalar stores wrote 1.24x the bytes of
// x and r at 512^3); the arithmetic and the order of the sums are those of
// the scalar form.
template <typename T>
__device__ __forceinline__ typename Vec2<T>::type ld_pair(const T* p) {
  return *reinterpret_cast<const typename Vec2<T>::type*>(p);
}
template <typename T>
__device__ __forceinline__ void st_pair(T* p, T a, T b) {
  typename Vec2<T>::type v;
  v.x = a;
  v.y = b;
  *reinterpret_cast<typename Vec2<T>::type*>(p) = v;
}

// r = b - A 0 = b, p = r, partial (r, r) (x taken as 0, :196-204)
template <typename T>
__global__ void __launch_bounds__(64 * CG_WAVES)
k_cg_start0(T* __restrict__ r, T* __restrict__ p, const T* __restrict__ b, Lvl<T> L, double* __restrict__ partial) {
  __shared__ double sh[CG_WAVES];
  double acc = cg_rows(L, [&](long q, bool two) {
    if (two) {
      const auto bv = ld_pair(b + q);
      const T r0 = bv.x - T(0), r1 = bv.y - T(0);
      st_pair(r + q, r0, r1);
      st_pair(p + q, r0, r1);
      return (double)r0 * (double)r0 + (double)r1 * (double)r1;
    }
    const T r0 = b[q] - T(0);
    r[q] = r0;
    p[q] = r0;
    return (double)r0 * (double)r0 + (double)T(0) * (double)T(0);
  });
  double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// x = x + alpha p, r = r - alpha Ap (:209-216), partial (r, r) (:221)
template <typename T>
__global__ void __launch_bounds__(64 * CG_WAVES)
k_cg_xr(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ ap,
        const double* __restrict__ sc, int rr_slot, Lvl<T> L, bool zero_x, double* __restrict__ partial) {
  __shared__ double sh[CG_WAVES];
  const T alpha = (T)(sc[rr_slot] / sc[2]);
  double acc = cg_rows(L, [&](long q, bool two) {
    if (two) {
      const auto pv = ld_pair(p + q), av = ld_pair(ap + q), rv = ld_pair(r + q);
      T x0 = T(0), x1 = T(0);
      if (!zero_x) {
        const auto xv = ld_pair(x + q);
        x0 = xv.x;
        x1 = xv.y;
      }
      const T n0 = x0 + alpha * pv.x, n1 = x1 + alpha * pv.y;
      const T s0 = rv.x - alpha * av.x, s1 = rv.y - alpha * av.y;
      st_pair(x + q, n0, n1);
      st_pair(r + q, s0, s1);
      double a = (double)s0 * (double)s0;
      a += (double)s1 * (double)s1;
      return a;
    }
    const T x0 = zero_x ? T(0) : x[q];
    const T n0 = x0 + alpha * p[q];
    const T s0 = r[q] - alpha * ap[q];
    x[q] = n0;
    r[q] = s0;
    return (double)s0 * (double)s0;
  });
  double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// p = r + beta p with beta = temp2 / temp1 (:223-229)
template <typename T>
__global__ void __launch_bounds__(64 * CG_WAVES)
k_cg_p(T* __restrict__ p, const T* __restrict__ r, const double* __restrict__ sc, int new_slot, int old_slot, Lvl<T> L) {
  const T beta = (T)(sc[new_slot] / sc[old_slot]);
  cg_rows(L, [&](long q, bool two) {
    if (two) {
      const auto rv = ld_pair(r + q), pv = ld_pair(p + q);
      st_pair(p + q, rv.x + beta * pv.x, rv.y + beta * pv.y);
    } else {
      p[q] = r[q] + beta * p[q];
    }
    return 0.0;
  });
}

// p2: a second direction buffer (with the march kernels the direction
// update of iteration it-1 is folded into iteration it's A p launch, which
// writes the new direction there; p and p2 then trade places)
template <typename T>
static void cg_iterate(Ops<T>& op, DevLevel<T>& d, T* x, const T* b, T* r, T* p, T* p2, T* ap, int nu, bool zero_x,
                       double* partial, double* sc) {
  const long nrow = (long)d.ny * d.nz;
  const int eb = (int)std::max(1L, std::min<long>((nrow + CG_WAVES - 1) / CG_WAVES, RED_BLOCKS));
  const dim3 eblk(64 * CG_WAVES);
  MarchGrid G = Ops<T>::march_grid(d, RESID_TY, false);
  const int mb = G.gx * G.gy * G.nch;
  const bool march = mb <= PARTIAL_CAP;
  // start: r = b - A x, p = r, temp1 = (r, r)  (:196-204)
  if (zero_x) {
    hipLaunchKernelGGL((k_cg_start0<T>), dim3(eb), eblk, 0, op.s, r, p, b, d.L(), partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, op.s, partial, eb, sc + 0);
  } else if (march) {
    hipLaunchKernelGGL((k_resid_march<T, RESID_TY, 4>), dim3(mb), dim3(MARCH_THREADS), 0, op.s, x, (const float*)nullptr,
                       b, r, (float*)nullptr, p, d.L(), G, partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, op.s, partial, mb, sc + 0);
  } else {
    op.resid_norm(d, x, b, r, partial, sc + 0);
    hipMemcpyAsync(p, r, d.ncell * sizeof(T), hipMemcpyDeviceToDevice, op.s);
  }
  if (nu <= 0) {
    if (zero_x) hipMemsetAsync(x, 0, d.ncell * sizeof(T), op.s);
    return;
  }
  for (int it = 0; it < nu; it++) {
    const int cur = it & 1, nxt = (it + 1) & 1;
    if (march && it > 0) {   // p = r + beta p (:223-229) folded into A p (:206-207)
      hipLaunchKernelGGL((k_cg_ap_fold<T, RESID_TY>), dim3(mb), dim3(MARCH_THREADS), 0, op.s, r, p, p2, ap, d.L(),
                         G, partial, sc, cur, nxt);
      hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, op.s, partial, mb, sc + 2);
      std::swap(p, p2);
    } else if (march) {                                             // :206-207
      hipLaunchKernelGGL((k_resid_march<T, RESID_TY, 3>), dim3(mb), dim3(MARCH_THREADS), 0, op.s, p,
                         (const float*)nullptr, (const T*)nullptr, ap, (float*)nullptr, (T*)nullptr, d.L(), G, partial);
      hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, op.s, partial, mb, sc + 2);
    } else {
      op.apply(d, ap, p);
      op.dot(d, ap, p, partial, sc + 2);
    }
    hipLaunchKernelGGL((k_cg_xr<T>), dim3(eb), eblk, 0, op.s, x, r, p, ap, sc, cur, d.L(), zero_x && it == 0,
                       partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, op.s, partial, eb, sc + nxt);   // :221
    if (!march && it + 1 < nu)
      hipLaunchKernelGGL((k_cg_p<T>), dim3(eb), eblk, 0, op.s, p, r, sc, nxt, cur, d.L());
  }
}

// ---------------------------------------------------------------------------
// Mixed-precision helpers
__global__ void k_f64_to_f32(float* __restrict__ dst, const double* __restrict__ src, long n) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) dst[c] = (float)src[c];
}
__global__ void k_f32_to_f64(double* __restrict__ dst, const float* __restrict__ src, long n) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) dst[c] = (double)src[c];
}

}  // namespace pepsmg

using namespace pepsmg;

// ---------------------------------------------------------------------------
int pepsmg_ctx::alloc_scratch() {
  HIP_TRY(hipMalloc(&d_partial, PARTIAL_CAP * sizeof(double)));
  HIP_TRY(hipMalloc(&d_scalars, 64 * sizeof(double)));
  HIP_TRY(hipMalloc(&d_flag, sizeof(int)));
  citers_cap = std::max(ncycle, 1) + 64;
  HIP_TRY(hipMalloc(&d_citers, citers_cap * sizeof(int)));
  HIP_TRY(hipEventCreate(&ev0));
  HIP_TRY(hipEventCreate(&ev1));
  HIP_TRY(hipEventCreateWithFlags(&ev_norm, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&ev_entry, hipEventDisableTiming));
  HIP_TRY(hipHostMalloc((void**)&h_scalars, 8 * sizeof(double), hipHostMallocDefault));
  return PEPSMG_OK;
}

void pepsmg_ctx::release() {
  cg64.release();
  cg32.release();
  h64.release();
  h32.release();
  hipFree(d_partial);
  hipFree(d_scalars);
  hipFree(d_flag);
  hipFree(d_citers);
  hipFree(d_pad);
  hipFree(d_pad2);
  hipFree(d_cinv);
  d_partial = d_scalars = d_pad = d_pad2 = d_cinv = nullptr;
  d_flag = d_citers = nullptr;
  if (ev0) hipEventDestroy(ev0);
  if (ev1) hipEventDestroy(ev1);
  if (ev_norm) hipEventDestroy(ev_norm);
  if (ev_entry) hipEventDestroy(ev_entry);
  ev0 = ev1 = ev_norm = ev_entry = nullptr;
  for (auto e : timer.ev)
    if (e) hipEventDestroy(e);
  timer.ev.clear();
  if (h_scalars) hipHostFree(h_scalars);
  h_scalars = nullptr;
  if (stream) hipStreamDestroy(stream);
  stream = nullptr;
}

namespace pepsmg {

// Measurement hook: time the plain fused launches on level d of this solve.
template <typename T>
static void timer_attach(pepsmg_ctx* c, Ops<T>& op, DevLevel<T>& d) {
  if (!c->timer.on) return;
  op.timer = &c->timer;
  c->timer.level = &d;
  c->timer.used = 0;
  op.fused_name(d, c->timer.name, sizeof c->timer.name);
}

// Adds the bracketed launches of the finished solve to the totals.
static void timer_collect(pepsmg_ctx* c) {
  KernelTimer& t = c->timer;
  for (int q = 0; q + 1 < t.used; q += 2) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, t.ev[q], t.ev[q + 1]) == hipSuccess) {
      t.total_ms += ms;
      t.launches++;
    }
  }
  t.used = 0;
  t.level = nullptr;
}

// Smoothing step on level l of hierarchy h (rbgs_iterator or cg_iterator
// on the level's x and b).  zero_start: x is 0 on entry (coarse levels on
// the way down, src/multigrid.f90:1937).  prolong_first: first add the
// prolongation of the next coarser level's x (src/multigrid.f90:1958-1959).
// last_mode: fused-kernel mode of the last iteration (4: it also stores x
// into the caller's array).
template <typename T>
static void smooth_level(pepsmg_ctx* c, Ops<T>& op, Hier<T>& h, CgBufs<T>* cg, int l, int nu, T omega,
                         bool zero_start, bool prolong_first, int last_mode = 0) {
  DevLevel<T>& d = h.lv[l];
  bool fused = c->opt.fused_smoother && c->opt.smoother == 0 && nu > 0;
  if (prolong_first) op.prolong_add(d, h.lv[l + 1], d.x);
  if (c->opt.smoother == 1) {
    cg_iterate(op, d, d.x, d.b, d.r, cg->p[l], cg->p2[l], cg->ap[l], nu, zero_start, c->d_partial,
               c->d_scalars + 8);
    return;
  }
  if (!fused) {
    op.rbgs(d, d.x, d.b, nu, omega, zero_start);
    return;
  }
  for (int it = 0; it < nu; it++)
    op.fused_iter(d, omega, (it == 0 && zero_start) ? 1 : (it == nu - 1 && last_mode ? last_mode : 0));
}

// Coarsest level: FP64 solve in one workgroup (src/multigrid.f90:1944-1949,
// tolerance = the V-cycle tolerance, at most 1000 iterations).
static void coarse_solve64(pepsmg_ctx* c, DevLevel<double>& d, double omega, double tol, int* slot) {
  hipLaunchKernelGGL((k_coarse_solve<double>), dim3(1), dim3(COARSE_THREADS), 0, c->stream, d.x, d.b, d.L(),
                     omega, tol, 1000, (int*)nullptr, slot);
}

template <typename T>
static void pc_launch(pepsmg_ctx* c, const PCArgs<T>& A, size_t lds) {
  set_dynamic_lds((const void*)k_coarse_vcycle<T>, c->lds_max);   // a failure surfaces at the solve's end
  hipLaunchKernelGGL((k_coarse_vcycle<T>), dim3(1), dim3(PC_THREADS), lds, c->stream, A);
}

// Coarsest level on the per-level path: FP64 solve in one workgroup.  In
// FP32 the right-hand side is widened to FP64 and the solution narrowed.
static void coarse_solve_level(pepsmg_ctx* c, Hier<double>& h, double omega, double tol, int* slot) {
  coarse_solve64(c, h.lv[c->nlevel], omega, tol, slot);
}
static void coarse_solve_level(pepsmg_ctx* c, Hier<float>& h, float omega, double tol, int* slot) {
  int L = c->nlevel;
  DevLevel<float>& cf = h.lv[L];
  DevLevel<double>& cd = c->h64.lv[L];
  long n = (long)cf.ncell;
  hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, cd.b, cf.b, n);
  coarse_solve64(c, cd, (double)omega, tol, slot);
  hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, cf.x, cd.x, n);
}

// The inverse of the coarsest level's operator (c0 on the diagonal, the
// neighbour coefficients beside it; ghost neighbours are 0), by Gauss-Jordan
// elimination with partial pivoting in FP64, stored transposed on the device
// (c->d_cinv).  The FP32 inner V-cycles of the mixed-precision solve apply it
// as their coarsest solve; levels larger than PC_CINV_MAX cells keep the
// red/black iterations.
static int ensure_coarse_inverse(pepsmg_ctx* c) {
  if (c->d_cinv) return PEPSMG_OK;
  const HostLevel& h = c->host[c->nlevel];
  const int nx = h.n[0], ny = h.n[1], nz = h.n[2], n = nx * ny * nz;
  if (n > PC_CINV_MAX) return PEPSMG_OK;
  std::vector<double> a((size_t)n * n, 0.0), inv((size_t)n * n, 0.0);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const int q = (i - 1) + nx * ((j - 1) + ny * (k - 1));
        const double cw = h.cw[0][i], ce = h.ce[0][i], cs = h.cw[1][j], cn = h.ce[1][j], cb = h.cw[2][k],
                     ct = h.ce[2][k];
        double* row = &a[(size_t)q * n];
        row[q] = -(((((cw + ce) + cs) + cn) + cb) + ct);
        if (i > 1) row[q - 1] = cw;
        if (i < nx) row[q + 1] = ce;
        if (j > 1) row[q - nx] = cs;
        if (j < ny) row[q + nx] = cn;
        if (k > 1) row[q - nx * ny] = cb;
        if (k < nz) row[q + nx * ny] = ct;
      }
  for (int q = 0; q < n; q++) inv[(size_t)q * n + q] = 1.0;
  for (int col = 0; col < n; col++) {
    int piv = col;
    for (int r = col + 1; r < n; r++)
      if (std::fabs(a[(size_t)r * n + col]) > std::fabs(a[(size_t)piv * n + col])) piv = r;
    if (!(std::fabs(a[(size_t)piv * n + col]) > 0.0)) return PEPSMG_OK;   // singular: keep the iterations
    if (piv != col)
      for (int k = 0; k < n; k++) {
        std::swap(a[(size_t)piv * n + k], a[(size_t)col * n + k]);
        std::swap(inv[(size_t)piv * n + k], inv[(size_t)col * n + k]);
      }
    const double d = a[(size_t)col * n + col];
    for (int k = 0; k < n; k++) {
      a[(size_t)col * n + k] /= d;
      inv[(size_t)col * n + k] /= d;
    }
    for (int r = 0; r < n; r++) {
      if (r == col) continue;
      const double f = a[(size_t)r * n + col];
      if (f == 0.0) continue;
      for (int k = 0; k < n; k++) {
        a[(size_t)r * n + k] -= f * a[(size_t)col * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)col * n + k];
      }
    }
  }
  std::vector<double> t((size_t)n * n);
  for (int l = 0; l < n; l++)
    for (int k = 0; k < n; k++) t[(size_t)k * n + l] = inv[(size_t)l * n + k];
  double* d = nullptr;
  HIP_TRY(hipMalloc(&d, t.size() * sizeof(double)));
  hipError_t e = hipMemcpy(d, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(d);
    return hip_fail(e, "hipMemcpy (coarsest inverse)");
  }
  c->d_cinv = d;
  return PEPSMG_OK;
}

// The down/up legs of one V-cycle below the fine level, for hierarchy h
// whose level-1 right-hand side has been restricted already
// (src/multigrid.f90:1936-1962).  Levels from the switch level s down run
// in the persistent single-workgroup kernel (RBGS smoother only); the
// coarsest solve is FP64 in every precision.  nu / nu_post: smoothing steps
// on the way down / up (equal in FP64, the reference's V(nu, nu)).
template <typename T>
static void vcycle_coarse_part(pepsmg_ctx* c, Ops<T>& op, Hier<T>& h, CgBufs<T>* cg, int nu, int nu_post, T omega,
                               double tol, int* slot) {
  int L = c->nlevel;
  int s = L + 1;
  PCArgs<T> A;
  size_t lds = 0;
  if (c->opt.smoother == 0 && c->opt.coarse_switch > 0) {
    for (int l = 1; l <= L; l++) {
      if (h.lv[l].cells() <= c->opt.coarse_switch &&
          pc_build(c->nlevel, c->h64, c->lds_max, h, l, nu, nu_post, omega, tol, slot, A, lds)) {
        s = l;
        break;
      }
    }
  }
  if constexpr (std::is_same<T, float>::value) A.cinv = c->d_cinv;   // exact coarsest solve (inner cycles)
  int last = std::min(s, L) - 1;   // last level of the per-level down-leg
  for (int l = 1; l <= last; l++) {
    DevLevel<T>& d = h.lv[l];
    smooth_level(c, op, h, cg, l, nu, omega, true, false);
    op.residual_restrict(d, h.lv[l + 1], d.x, d.b);
  }
  if (s <= L) {
    A.gxs = h.lv[s].x;   // current buffers
    A.gbs = h.lv[s].b;
    pc_launch(c, A, lds);
  } else {
    coarse_solve_level(c, h, omega, tol, slot);
  }
  for (int l = last; l >= 1; l--) smooth_level(c, op, h, cg, l, nu_post, omega, false, true);
}

static void print_cycle(int cyc, double rv, double r0) {
  std::printf("[MG] cycle = %4d, Error: %15.7E %15.7E %15.7E\n", cyc, std::sqrt(rv), std::sqrt(r0),
              r0 > 0.0 ? std::sqrt(rv / r0) : 0.0);
}

// The initial residual norm |b - A x0|^2.  Not finite: rhs or sol holds a
// NaN or an infinity (the reference would cycle on NaN until its cap).
// Zero: x0 solves the system exactly (the reference would form sqrt(0/0)
// and run every cycle); the solve then stops with rho = 0.
static int check_res0(double res0) {
  if (!std::isfinite(res0))
    return fail(PEPSMG_ERR_ARG, "|b - A x0|^2 = %g is not finite: rhs or sol holds NaN or Inf", res0);
  return PEPSMG_OK;
}

// rho of a cycle's residual |b - A x|^2 = r2 against res0 (0 when res0 is 0)
static double rel_residual(double r2, double res0) { return res0 > 0.0 ? std::sqrt(r2 / res0) : 0.0; }

static int diverged(int cyc, double rho) {
  return fail(PEPSMG_ERR_DIVERGED, "the relative residual became %g after %d cycle(s)", rho, cyc);
}

// FP64 V-cycles (src/multigrid.f90:1907-1976).  ux, ub: the caller's
// padded sol and rhs (device), read directly by the first smoothing
// iteration (fused kernel mode 3), which stores b in the solver's layout;
// the last iteration of a cycle expected to end the solve also stores x into
// ux (mode 4), and c->sol_out_done records whether the solve ended on such a
// cycle.  nullptr (or a path without the fused smoother's norm fold): x and
// b are already in h64.lv[0].x / .b.
static int solve_fp64(pepsmg_ctx* c, int nu, double tol, double omega, pepsmg_stats* st, std::vector<double>& hist,
                      std::vector<int>& cit, double* res0_out, int* ncyc_out, double* rho_out, double* ux = nullptr,
                      const double* ub = nullptr) {
  Hier<double>& h = c->h64;
  Ops<double> op{c->stream};
  op.tile_y = c->opt.tile_y;
  op.prefetch = c->opt.fused_smoother == 2 ? 2 : 1;
  op.tile_x = c->opt.tile_x;
  timer_attach(c, op, h.lv[0]);
  CgBufs<double>* cg = nullptr;
  if (c->opt.smoother == 1) {
    cg = &c->cg64;
    int rc = cg->ensure(h);
    if (rc) return rc;
  }
  DevLevel<double>& f = h.lv[0];
  double hs[2];
  // With the fused RBGS smoother the convergence norm of a cycle's result
  // (:1970-1971) comes out of the first pre-smoothing iteration of the next
  // cycle, which evaluates |b - A x|^2 of the x it starts from while it
  // smooths; when the tolerance is met that iteration's output is dropped.
  // The initial norm (:1924-1925) comes out of cycle 1 the same way, and a
  // separate pass is left only after the last permitted cycle.
  const bool fold = c->opt.fused_smoother && c->opt.smoother == 0 && nu > 0 && c->ncycle > 0;
  const int nbn = fold ? op.fused_blocks(f) : 0;
  const bool direct = fold && ux != nullptr;
  c->sol_out_done = false;
  if (direct) op.set_caller(f, ux, ub);
  int dual_cycle = -1;   // the cycle whose last iteration also stored x into ux
  double res0 = 0.0;
  int ncyc = 0;
  double rho = 1.0;
  int first = 1;   // first cycle to run
  if (!fold) {
    op.resid_norm_march(f, f.x, f.b, nullptr, c->d_partial, c->d_scalars + 0);  // :1924-1925
    HIP_TRY(hipMemcpyAsync(hs, c->d_scalars, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    res0 = hs[0];
    int rc = check_res0(res0);
    if (rc) return rc;
    if (res0 == 0.0) {   // x0 is exact: no cycle
      rho = 0.0;
      first = c->ncycle + 1;
    }
  }
  // the per-cycle factor of the last two cycles predicts that the next
  // check meets the tolerance (factor-2 margin)
  auto likely_converged = [&]() {
    const size_t n = hist.size();
    if (n < 2 || !(hist[n - 2] > 0.0)) return false;
    return hist[n - 1] * (hist[n - 1] / hist[n - 2]) < 2.0 * tol;
  };
  // :1971-1973; 1: rho < tol (stop), 0: go on, < 0: error (a non-finite
  // initial norm, or a cycle whose norm is no longer finite)
  auto check = [&](int cyc, double r2) -> int {
    int rc = check_res0(res0);
    if (rc) return rc;
    rho = rel_residual(r2, res0);
    hist.push_back(rho);
    ncyc = cyc;
    if (c->opt.verbose) print_cycle(cyc, r2, res0);
    if (!std::isfinite(rho)) return diverged(cyc, rho);
    return rho < tol ? 1 : 0;
  };
  // the cycle is predicted (the last two cycles' factor, no margin: a wrong
  // prediction costs the extra stores, a missed one the copy-out pass, about
  // the same) or bound (cycle cap) to end the solve; never before the first
  // read-back, which brings the ghost flag of sol (nothing is stored into an
  // array the entry point will reject)
  auto ends_solve = [&](int cyc) {
    if (!direct || cyc < 2 || *c->ghost_flag_host() != 0) return false;
    if (cyc == c->ncycle) return true;
    const size_t n = hist.size();
    if (n < 2 || !(hist[n - 2] > 0.0)) return false;
    return hist[n - 1] * (hist[n - 1] / hist[n - 2]) < tol;
  };
  for (int cyc = first; cyc <= c->ncycle; cyc++) {
    if (fold) {
      // first pre-smoothing iteration, with |b - A x|^2 of the incoming x (in
      // cycle 1 of a direct solve from the caller's arrays)
      op.fused_iter(f, omega, (cyc == 1 && direct) ? 3 : 2, c->d_partial);
      hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, c->stream, c->d_partial, nbn,
                         c->d_scalars + (cyc == 1 ? 0 : 1));
      int pre_done = 1;
      if (cyc > 1) {
        HIP_TRY(hipMemcpyAsync(c->h_scalars, c->d_scalars, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipEventRecord(c->ev_norm, c->stream));
        // While the host waits for the norm, the GPU runs the second
        // iteration into the spare buffer f.r; the first iteration's input,
        // the answer if the tolerance is met, stays in f.x2.  Skipped when
        // the check is predicted to stop the solve.
        const bool ahead = nu >= 2 && !likely_converged();
        if (ahead) op.fused_iter_into(f, omega, f.r);
        HIP_TRY(hipEventSynchronize(c->ev_norm));
        res0 = c->h_scalars[0];
        int stop = check(cyc - 1, c->h_scalars[1]);
        if (stop < 0) return stop;
        if (stop) {
          std::swap(f.x, f.x2);   // the solution is the iteration's input
          c->sol_out_done = dual_cycle == cyc - 1;
          break;
        }
        if (ahead) {
          std::swap(f.x, f.r);   // the iterate after two iterations; f.r is spare again
          pre_done = 2;
        }
      }
      smooth_level(c, op, h, cg, 0, nu - pre_done, omega, false, false);  // :1931 (remaining iterations)
    } else {
      smooth_level(c, op, h, cg, 0, nu, omega, false, false);            // :1931
    }
    op.residual_restrict(f, h.lv[1], f.x, f.b);                            // :1932-1933
    int* slot = (cyc - 1 < c->citers_cap) ? c->d_citers + (cyc - 1) : nullptr;
    vcycle_coarse_part<double>(c, op, h, cg, nu, nu, omega, tol, slot);    // :1936-1962
    const bool dual = ends_solve(cyc);
    if (dual) dual_cycle = cyc;
    smooth_level(c, op, h, cg, 0, nu, omega, false, true, dual ? 4 : 0);  // :1965-1967
    if (!fold || cyc == c->ncycle) {
      c->sol_out_done = dual_cycle == cyc;
      op.resid_norm_march(f, f.x, f.b, nullptr, c->d_partial, c->d_scalars + 1);  // :1970-1971
      HIP_TRY(hipMemcpyAsync(hs, c->d_scalars, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      if (fold) res0 = hs[0];
      int stop = check(cyc, hs[1]);
      if (stop < 0) return stop;
      if (stop) break;
    }
  }
  *res0_out = res0;
  if (ncyc > 0) {
    int m = std::min(ncyc, c->citers_cap);
    cit.resize(m);
    HIP_TRY(hipMemcpy(cit.data(), c->d_citers, m * sizeof(int), hipMemcpyDeviceToHost));
  }
  *ncyc_out = ncyc;
  *rho_out = rho;
  (void)st;
  return PEPSMG_OK;
}

// FP32 V-cycles inside FP64 iterative refinement.  The fine residual
// b - A x, the update x += e and the coarsest solve stay in FP64.
// ux, ub: the caller's padded x and b (device).  b is read from there by
// every refinement step; x until the first update, which stores x + e in
// the solver's layout.  The step predicted to be the last stores its x
// into ux.  c->sol_out_done records whether ux holds the solution when the
// solve ends (no update at all, or a last step that stored into it).
// nullptr: x and b are already in the solver's layout.
static int solve_mixed(pepsmg_ctx* c, int nu, double tol, double omega, std::vector<double>& hist,
                       std::vector<int>& cit, double* res0_out, int* ncyc_out, int* nref_out, double* rho_out,
                       double* ux, const double* ub) {
  Hier<double>& h = c->h64;
  Hier<float>& g = c->h32;
  Ops<float> op{c->stream};
  op.tile_y = c->opt.tile_y;
  op.prefetch = c->opt.fused_smoother == 2 ? 2 : 1;
  op.tile_x = c->opt.tile_x;
  timer_attach(c, op, g.lv[0]);
  CgBufs<float>* cg = nullptr;
  if (c->opt.smoother == 1) {
    cg = &c->cg32;
    int rc = cg->ensure(g);
    if (rc) return rc;
  }
  {
    int rc = ensure_coarse_inverse(c);
    if (rc) return rc;
  }
  DevLevel<double>& f = h.lv[0];
  DevLevel<float>& f32 = g.lv[0];
  const bool big = f.ny >= 512;   // refinement kernels' tile height RESID_TY_BIG, else RESID_TY
  MarchGrid G = Ops<double>::march_grid(f, big ? RESID_TY_BIG : RESID_TY, false);
  int nb = G.gx * G.gy * G.nch;
  // the caller's padded layout (pitches nx+2, (nx+2)(ny+2); offset of cell (1,1,1))
  Lvl<double> U = f.L();
  U.px = f.nx + 2;
  U.pxy = (long)(f.nx + 2) * (f.ny + 2);
  const long uo = 1 + U.px + U.pxy;
  // where x and b are read from: the caller's arrays, or the solver's layout
  const Lvl<double> LB = ux ? U : f.L();
  const double* bsrc = ux ? ub + uo : f.b;
  bool x_caller = false;   // x lives in the caller's array only (ux)
  c->sol_out_done = false;
  // r32 = b - A x and its norm from the caller's arrays; x stays there
  auto start_from_caller = [&]() {
    if (big)
      hipLaunchKernelGGL((k_refine_start<RESID_TY_BIG>), dim3((unsigned)nb), dim3(MARCH_THREADS), 0, c->stream,
                         ux + uo, ub + uo, f32.b, f.L(), U, G, c->d_partial);
    else
      hipLaunchKernelGGL((k_refine_start<RESID_TY>), dim3((unsigned)nb), dim3(MARCH_THREADS), 0, c->stream, ux + uo,
                         ub + uo, f32.b, f.L(), U, G, c->d_partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, c->stream, c->d_partial, nb, c->d_scalars);
    x_caller = true;
    c->sol_out_done = true;
  };
  // x2 = x + e; r32 = float(b - A x2); partial |b - A x2|^2; then swap x,
  // x2 (x read from the caller's array right after a start).  last: x + e
  // goes only into the caller's array, with its norm; if the step then
  // misses the tolerance, start_from_caller forms r32 from there.
  auto refine_update = [&](bool last) {
    if (last && big)
      hipLaunchKernelGGL((k_refine_last<RESID_TY_BIG>), dim3((unsigned)nb), dim3(MARCH_THREADS), 0, c->stream,
                         f.x, f32.x, bsrc, f.L(), LB, U, G, c->d_partial, ux + uo);
    else if (last)
      hipLaunchKernelGGL((k_refine_last<RESID_TY>), dim3((unsigned)nb), dim3(MARCH_THREADS), 0, c->stream, f.x,
                         f32.x, bsrc, f.L(), LB, U, G, c->d_partial, ux + uo);
    else {
      const double* xs = x_caller ? ux + uo : f.x;
      const Lvl<double> LX = x_caller ? U : f.L();
      if (big)
        hipLaunchKernelGGL((k_refine_step<RESID_TY_BIG>), dim3((unsigned)nb), dim3(MARCH_THREADS), 0, c->stream, xs,
                           f32.x, bsrc, f32.b, f.x2, f.L(), LX, LB, G, c->d_partial);
      else
        hipLaunchKernelGGL((k_refine_step<RESID_TY>), dim3((unsigned)nb), dim3(MARCH_THREADS), 0, c->stream, xs,
                           f32.x, bsrc, f32.b, f.x2, f.L(), LX, LB, G, c->d_partial);
    }
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, c->stream, c->d_partial, nb, c->d_scalars);
    if (!last) {
      std::swap(f.x, f.x2);
      x_caller = false;
    }
    c->sol_out_done = last;
  };
  // r32 = b - A x and res0, with e = 0
  if (ux) {   // x and b are read from the caller's padded arrays
    start_from_caller();
  } else {
    hipMemsetAsync(f32.x, 0, f32.ncell * sizeof(float), c->stream);
    refine_update(false);
  }
  // SOR factor of the FP32 inner cycles (an inner-solver setting: the FP64
  // residual and update, which decide convergence, do not use it)
  float om = (float)(c->opt.inner_omega > 0.0 ? c->opt.inner_omega : omega);
  // pre- and post-smoothing steps of the FP32 V-cycles (the FP64 residual
  // and update do not depend on them): inner_nu (or the caller's
  // maxiteration) and inner_nu_post (or the pre-smoothing count)
  const int nui = c->opt.inner_nu > 0 ? c->opt.inner_nu : nu;
  const int nup = c->opt.inner_nu_post > 0 ? c->opt.inner_nu_post : nui;
  // The host reads each norm while the GPU already runs the next step's
  // first pre-smoothing (it needs only r32, and touches no FP64 array), so
  // the read-back leaves no gap; skipped when the step is predicted to stop.
  bool ahead = false;
  auto read_norm = [&](bool run_ahead) -> int {
    HIP_TRY(hipMemcpyAsync(c->h_scalars, c->d_scalars, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(c->ev_norm, c->stream));
    ahead = run_ahead;
    if (ahead) smooth_level(c, op, g, cg, 0, nui, om, true, false);
    HIP_TRY(hipEventSynchronize(c->ev_norm));
    return PEPSMG_OK;
  };
  int rc = read_norm(c->ncycle > 0 && c->opt.max_refine > 0);
  if (rc) return rc;
  double res0 = c->h_scalars[0];
  *res0_out = res0;
  rc = check_res0(res0);
  if (rc) return rc;
  int ncyc = 0, nref = 0;
  double rho = res0 == 0.0 ? 0.0 : 1.0;   // x0 exact: no refinement step
  // Per-V-cycle reduction factor of the previous step.  A step that is
  // predicted (with a factor-2 margin) to reach the tolerance in fewer than
  // inner_cycles cycles runs only those: typically the last step, which needs
  // a small reduction, while FP32 storage caps every step at about 1e-3 (so
  // a capped step overstates the factor and the prediction stays safe).
  double fcyc = 0.0;
  for (int s = 1; s <= c->opt.max_refine && ncyc < c->ncycle && res0 > 0.0; s++) {
    int nin = c->opt.inner_cycles;
    if (fcyc > 0.0 && fcyc < 1.0) {
      double need = 0.5 * tol / rho, red = fcyc;
      int k = 1;
      while (k < nin && red > need) {
        k++;
        red *= fcyc;
      }
      nin = k;
    }
    nin = std::min(nin, c->ncycle - ncyc);   // number_of_vcycles caps the FP32 V-cycles of the solve
    double rho_prev = rho;
    for (int q = 0; q < nin; q++) {
      int* slot = (ncyc < c->citers_cap) ? c->d_citers + ncyc : nullptr;
      if (!(q == 0 && ahead)) smooth_level(c, op, g, cg, 0, nui, om, q == 0, false);
      op.residual_restrict(f32, g.lv[1], f32.x, f32.b);
      vcycle_coarse_part<float>(c, op, g, cg, nui, nup, om, tol, slot);
      smooth_level(c, op, g, cg, 0, nup, om, false, true);
      ncyc++;
    }
    // predicted to stop here: the previous per-cycle factor over this step's cycles, factor-2 margin
    const bool stop_likely = fcyc > 0.0 && rho_prev * std::pow(fcyc, nin) < 2.0 * tol;
    // (never into an array the entry point will reject: its ghost flag has
    // arrived with the first norm; never while x is read from that array)
    const bool last = stop_likely && ux != nullptr && !x_caller && *c->ghost_flag_host() == 0;
    refine_update(last);
    rc = read_norm(!stop_likely && s < c->opt.max_refine && ncyc < c->ncycle);
    if (rc) return rc;
    const double r2 = c->h_scalars[0];
    rho = rel_residual(r2, res0);
    hist.push_back(rho);
    nref = s;
    if (c->opt.verbose) print_cycle(s, r2, res0);
    if (!std::isfinite(rho)) {
      // (a step predicted to be the last has written its x + e into sol)
      *ncyc_out = ncyc;
      *nref_out = nref;
      *rho_out = rho;
      return diverged(ncyc, rho);
    }
    if (rho < tol) break;
    if (last) start_from_caller();   // mispredicted: x + e is in the caller's array only
    if (rho > 0.0 && rho_prev > 0.0) fcyc = std::pow(rho / rho_prev, 1.0 / nin);
  }
  if (ncyc > 0) {
    int m = std::min(ncyc, c->citers_cap);
    cit.resize(m);
    HIP_TRY(hipMemcpy(cit.data(), c->d_citers, m * sizeof(int), hipMemcpyDeviceToHost));
  }
  *ncyc_out = ncyc;
  *nref_out = nref;
  *rho_out = rho;
  return PEPSMG_OK;
}

template <typename T>
static void pack(hipStream_t s, T* dst, const double* src, const DevLevel<T>& d, int* flag) {
  hipLaunchKernelGGL((k_pack<T>), grid3(d.nx + 2, d.ny + 2, d.nz + 2, 64, 4), dim3(64, 4, 1), 0, s, dst, src, d.L(),
                     flag);
}
template <typename T>
static void unpack(hipStream_t s, double* dst, const T* src, const DevLevel<T>& d, int zero_ghosts) {
  hipLaunchKernelGGL((k_unpack<T>), grid3(d.nx + 2, d.ny + 2, d.nz + 2, 64, 4), dim3(64, 4, 1), 0, s, dst, src,
                     d.L(), zero_ghosts);
}

// Compare a reference AoS matrix with the metric rediscretisation.
static int check_matrix(const HostLevel& h, const double* coeff) {
  if (!coeff) return PEPSMG_OK;
  int nx = h.n[0], ny = h.n[1], nz = h.n[2];
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const double* q = coeff + 7 * ((size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k - 1)));
        double c1 = h.cw[0][i], c2 = h.ce[0][i], c3 = h.cw[1][j], c4 = h.ce[1][j], c5 = h.cw[2][k], c6 = h.ce[2][k];
        double c0 = -(c1 + c2 + c3 + c4 + c5 + c6);
        if (q[0] != c0 || q[1] != c1 || q[2] != c2 || q[3] != c3 || q[4] != c4 || q[5] != c5 || q[6] != c6)
          return fail(PEPSMG_ERR_MATRIX,
                      "coefficient matrix differs from the rediscretisation of the grid metrics at (%d,%d,%d)", i, j, k);
      }
  return PEPSMG_OK;
}

static int check_ghosts_host(const double* a, int nx, int ny, int nz, const char* name) {
  size_t sx = nx + 2, sy = ny + 2;
  for (int k = 0; k <= nz + 1; k++)
    for (int j = 0; j <= ny + 1; j++) {
      bool full = (k == 0 || k == nz + 1 || j == 0 || j == ny + 1);
      const double* row = a + sx * ((size_t)j + sy * k);
      if (full) {
        for (int i = 0; i <= nx + 1; i++)
          if (row[i] != 0.0) return fail(PEPSMG_ERR_GHOST, "%s has a non-zero ghost cell at (%d,%d,%d)", name, i, j, k);
      } else if (row[0] != 0.0 || row[nx + 1] != 0.0) {
        return fail(PEPSMG_ERR_GHOST, "%s has a non-zero ghost cell in row (%d,%d)", name, j, k);
      }
    }
  return PEPSMG_OK;
}

static int run_solve(pepsmg_ctx* c, int nu, double tol, double omega, pepsmg_stats* st, double* ux = nullptr,
                     const double* ub = nullptr) {
  std::vector<double> hist;
  std::vector<int> cit;
  double res0 = 0, rho = 0;
  int ncyc = 0, nref = 0, rc;
  (void)hipGetLastError();   // an error the caller's earlier HIP calls left is not this solve's
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  if (c->opt.precision == 1)
    rc = solve_mixed(c, nu, tol, omega, hist, cit, &res0, &ncyc, &nref, &rho, ux, ub);
  else
    rc = solve_fp64(c, nu, tol, omega, st, hist, cit, &res0, &ncyc, &rho, ux, ub);
  if (rc && rc != PEPSMG_ERR_DIVERGED) return rc;   // (a diverged solve still reports its statistics)
  const int solve_rc = rc;
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev1));
  if (c->timer.on) timer_collect(c);
  HIP_TRY(hipGetLastError());   // a launch the runtime rejected leaves its error here
  if (st) {
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    st->device_ms = ms;
    st->cycles = ncyc;
    st->refine_steps = nref;
    st->res0 = res0;
    st->rho = rho;
    st->converged = rho < tol ? 1 : 0;
    if (st->history)
      for (int q = 0; q < (int)hist.size() && q < st->history_cap; q++) st->history[q] = hist[q];
    if (st->coarse_iters)
      for (int q = 0; q < (int)cit.size() && q < st->history_cap; q++) st->coarse_iters[q] = cit[q];
  }
  return solve_rc;
}

}  // namespace pepsmg

extern "C" {

int pepsmg_multigrid_solve_vcycle_device(pepsmg_ctx* c, double* d_sol, double* d_rsd, const double* d_rhs,
                                         int32_t maxiteration, double tolerance, double omega_sor, void* stream,
                                         pepsmg_stats* st) {
  if (!c || !d_sol || !d_rhs) return fail(PEPSMG_ERR_ARG, "NULL argument");
  if (maxiteration < 0) return fail(PEPSMG_ERR_ARG, "maxiteration < 0");
  hipStream_t saved = c->stream;
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    // the context's own stream does not synchronize with the default
    // stream: order it after the work the caller queued there (for example
    // the writes of sol and rhs)
    HIP_TRY(hipEventRecord(c->ev_entry, (hipStream_t)0));
    HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_entry, 0));
  }
  DevLevel<double>& f = c->h64.lv[0];
  int rc = PEPSMG_OK;
  int flag = 0;
  // Mixed precision, and FP64 with the fused smoother's norm fold: the solve
  // reads the caller's arrays itself (FP64: the first smoothing iteration,
  // which leaves x and b in the solver's layout; mixed precision: every
  // refinement step reads b there, and x until the first update), so only
  // the ghost layer of sol is checked here; other FP64 paths: copy-in passes
  // (with the same check).
  const bool direct = c->opt.precision == 1 || (c->opt.fused_smoother && c->opt.smoother == 0 && maxiteration > 0 &&
                                                c->ncycle > 0);
  hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream);
  if (direct) {
    const long items = 2L * (f.nx + 2) * (f.ny + 2) + 2L * f.nz * (f.nx + 2) + (long)f.ny * f.nz;
    const unsigned nb = (unsigned)std::min<long>((items + 255) / 256, 2048);
    hipLaunchKernelGGL(k_ghost_check, dim3(nb), dim3(256), 0, c->stream, d_sol, f.nx, f.ny, f.nz, c->d_flag);
  } else {
    pack(c->stream, f.x, d_sol, f, c->d_flag);
    pack(c->stream, f.b, d_rhs, f, (int*)nullptr);
  }
  // the flag is read back with the solve's first norm (the solve never
  // reads ghost cells and writes nothing of the caller's before the check)
  int* h_flag = c->ghost_flag_host();
  *h_flag = 0;
  hipError_t e = hipMemcpyAsync(h_flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) rc = hip_fail(e, "pack");
  if (!rc) rc = run_solve(c, maxiteration, tolerance, omega_sor, st, direct ? d_sol : nullptr, direct ? d_rhs : nullptr);
  if (rc) {
    // nothing of this solve may still run once the caller has the error
    (void)hipStreamSynchronize(c->stream);
  } else {
    e = hipStreamSynchronize(c->stream);
    flag = *h_flag;
    if (e != hipSuccess) rc = hip_fail(e, "pack");
    else if (flag) rc = fail(PEPSMG_ERR_GHOST, "sol has non-zero ghost cells (homogeneous Dirichlet required)");
  }
  if (!rc) {
    if (!(direct && c->sol_out_done)) unpack(c->stream, d_sol, f.x, f, 0);   // else the last step stored it
    if (d_rsd) {
      // mixed precision: b was only read from rhs, and x is in sol when the
      // solve ended there (the FP64 path stored x and b into both)
      if (c->opt.precision == 1) {
        pack(c->stream, f.b, d_rhs, f, (int*)nullptr);
        if (c->sol_out_done) pack(c->stream, f.x, d_sol, f, (int*)nullptr);
      }
      Ops<double> op{c->stream};
      op.residual(f, f.r, f.x, f.b);                                   // :1970
      unpack(c->stream, d_rsd, f.r, f, 1);
    }
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(e, "unpack");
  }
  c->stream = saved;
  return rc;
}

int pepsmg_multigrid_solve_vcycle(pepsmg_ctx* c, double* sol, double* rsd, const double* coeff, const double* rhs,
                                  int32_t maxiteration, double tolerance, double omega_sor, pepsmg_stats* st) {
  if (!c || !sol || !rhs) return fail(PEPSMG_ERR_ARG, "NULL argument");
  const HostLevel& h = c->host[0];
  int rc = check_matrix(h, coeff);
  if (rc) return rc;
  rc = check_ghosts_host(sol, h.n[0], h.n[1], h.n[2], "sol");
  if (rc) return rc;
  size_t np = (size_t)(h.n[0] + 2) * (h.n[1] + 2) * (h.n[2] + 2);
  if (c->pad_elems < np) {
    hipFree(c->d_pad);
    hipFree(c->d_pad2);
    c->d_pad = c->d_pad2 = nullptr;
    HIP_TRY(hipMalloc(&c->d_pad, np * sizeof(double)));
    HIP_TRY(hipMalloc(&c->d_pad2, np * sizeof(double)));
    c->pad_elems = np;
  }
  HIP_TRY(hipMemcpy(c->d_pad, sol, np * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_pad2, rhs, np * sizeof(double), hipMemcpyHostToDevice));
  // the rhs staging receives rsd after the solve (b is copied into the solver's layout before rsd is formed)
  rc = pepsmg_multigrid_solve_vcycle_device(c, c->d_pad, rsd ? c->d_pad2 : nullptr, c->d_pad2, maxiteration,
                                            tolerance, omega_sor, nullptr, st);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(sol, c->d_pad, np * sizeof(double), hipMemcpyDeviceToHost));
  if (rsd) HIP_TRY(hipMemcpy(rsd, c->d_pad2, np * sizeof(double), hipMemcpyDeviceToHost));
  return PEPSMG_OK;
}

int pepsmg_bench_smoother(pepsmg_ctx* c, int32_t precision, int32_t reps, double omega, double* avg_ms,
                          double* bytes, const char** kernel_name) {
  if (!c || reps < 1) return fail(PEPSMG_ERR_ARG, "invalid argument");
  if (precision == 1 && c->h32.lv.empty()) return fail(PEPSMG_ERR_ARG, "FP32 hierarchy not built");
  static thread_local char name[96];
  float ms = 0;
  double n = 0;
  const bool fused = c->opt.fused_smoother != 0;
  // fused: one red+black iteration reads x and b and writes x once (3 n
  // elements); colour sweep: reads x and b, writes half of x (2 n)
  auto run = [&](auto& op, auto& d, const char* tname) -> int {
    op.tile_y = c->opt.tile_y;
    op.tile_x = c->opt.tile_x;
    op.prefetch = c->opt.fused_smoother == 2 ? 2 : 1;
    using T = typename std::remove_reference<decltype(d.x[0])>::type;
    auto body = [&]() {
      for (int q = 0; q < reps; q++) {
        if (fused) op.fused_iter(d, (T)omega, 0);
        else op.smooth_colour(d, d.b, d.x, (T)omega, q & 1, false);
      }
    };
    body();
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    body();
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    n = (double)d.cells();
    if (fused)
      std::snprintf(name, sizeof name, "k_rb_fused<%s, %d, %d, 0, %d>", tname, op.tx_used(d), op.ty_used(d),
                    op.ty_used(d) == 32 ? 1 : op.prefetch);   // 32-row tiles load one plane ahead
    else
      std::snprintf(name, sizeof name, "k_rb_colour<%s, false>", tname);
    return PEPSMG_OK;
  };
  int rc;
  if (precision == 1) {
    Ops<float> op{c->stream};
    rc = run(op, c->h32.lv[0], "float");
  } else {
    Ops<double> op{c->stream};
    rc = run(op, c->h64.lv[0], "double");
  }
  if (rc) return rc;
  HIP_TRY(hipEventSynchronize(c->ev1));
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (avg_ms) *avg_ms = ms / reps;
  if (bytes) *bytes = (fused ? 3.0 : 2.0) * n * (precision == 1 ? 4.0 : 8.0);
  if (kernel_name) *kernel_name = name;
  return PEPSMG_OK;
}

int pepsmg_kernel_timing(pepsmg_ctx* c, int32_t enable) {
  if (!c) return fail(PEPSMG_ERR_ARG, "NULL context");
  KernelTimer& t = c->timer;
  if (enable && t.ev.empty()) {
    constexpr int POOL = 512;   // 256 timed launches per solve
    t.ev.resize(POOL, nullptr);
    for (auto& e : t.ev) HIP_TRY(hipEventCreate(&e));
  }
  t.on = enable != 0;
  t.used = 0;
  if (t.on) {   // switching off keeps the totals for pepsmg_kernel_timing_read
    t.total_ms = 0.0;
    t.launches = 0;
    t.name[0] = 0;
  }
  return PEPSMG_OK;
}

int pepsmg_kernel_timing_read(const pepsmg_ctx* c, double* total_ms, int32_t* launches, const char** kernel_name) {
  if (!c) return fail(PEPSMG_ERR_ARG, "NULL context");
  if (total_ms) *total_ms = c->timer.total_ms;
  if (launches) *launches = (int32_t)c->timer.launches;
  if (kernel_name) *kernel_name = c->timer.name;
  return PEPSMG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Standalone operators on one grid (host arrays).
namespace pepsmg {

// One grid on the device for the standalone operators: metrics, the level's
// arrays in element type T (coefficients converted from the FP64 metrics),
// and padded FP64 staging buffers for the caller's arrays.
template <typename T>
struct OneLevelT {
  HostLevel h;
  DevLevel<T> d;
  double *pad_a = nullptr, *pad_b = nullptr, *partial = nullptr, *sc = nullptr;
  int* flag = nullptr;
  hipStream_t s = nullptr;
  size_t np = 0;
  ~OneLevelT() {
    hipFree(d.cw[0]); hipFree(d.cw[1]); hipFree(d.cw[2]);
    hipFree(d.ce[0]); hipFree(d.ce[1]); hipFree(d.ce[2]);
    hipFree(d.x); hipFree(d.b); hipFree(d.r); hipFree(d.x2);
    hipFree(pad_a); hipFree(pad_b); hipFree(partial); hipFree(sc); hipFree(flag);
    if (s) hipStreamDestroy(s);
  }
  int init(const pepsmg_grid* g) {
    if (int rc = check_grid(g)) return rc;
    h.n[0] = g->nx; h.n[1] = g->ny; h.n[2] = g->nz;
    const double* dms[3] = {g->dxm, g->dym, g->dzm};
    const double* dgs[3] = {g->dxg, g->dyg, g->dzg};
    for (int a = 0; a < 3; a++) {
      h.dm[a].assign(dms[a], dms[a] + h.n[a] + 2);
      h.dg[a].assign(dgs[a], dgs[a] + h.n[a] + 2);
    }
    (void)hipGetLastError();   // an error the caller's earlier HIP calls left is not this call's
    h.setup_coeffs();
    d.nx = h.n[0]; d.ny = h.n[1]; d.nz = h.n[2];
    d.px = pitch_for(d.nx);
    d.pxy = d.px * d.ny;
    d.ncell = (size_t)d.pxy * d.nz;
    for (int a = 0; a < 3; a++) {
      const std::vector<T> cw(h.cw[a].begin(), h.cw[a].end()), ce(h.ce[a].begin(), h.ce[a].end());
      HIP_TRY(hipMalloc(&d.cw[a], cw.size() * sizeof(T)));
      HIP_TRY(hipMalloc(&d.ce[a], ce.size() * sizeof(T)));
      HIP_TRY(hipMemcpy(d.cw[a], cw.data(), cw.size() * sizeof(T), hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d.ce[a], ce.data(), ce.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMalloc(&d.x, d.ncell * sizeof(T)));
    HIP_TRY(hipMalloc(&d.b, d.ncell * sizeof(T)));
    HIP_TRY(hipMalloc(&d.r, d.ncell * sizeof(T)));
    HIP_TRY(hipMalloc(&d.x2, d.ncell * sizeof(T)));
    np = (size_t)(d.nx + 2) * (d.ny + 2) * (d.nz + 2);
    HIP_TRY(hipMalloc(&pad_a, np * sizeof(double)));
    HIP_TRY(hipMalloc(&pad_b, np * sizeof(double)));
    HIP_TRY(hipMalloc(&partial, PARTIAL_CAP * sizeof(double)));
    HIP_TRY(hipMalloc(&sc, 16 * sizeof(double)));
    HIP_TRY(hipMalloc(&flag, sizeof(int)));
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return PEPSMG_OK;
  }
  int upload(T* dst, const double* host) {
    HIP_TRY(hipMemcpy(pad_a, host, np * sizeof(double), hipMemcpyHostToDevice));
    pack(s, dst, pad_a, d, (int*)nullptr);
    HIP_TRY(hipStreamSynchronize(s));
    return PEPSMG_OK;
  }
  int download(double* host, const T* src, int zero_ghosts) {
    HIP_TRY(hipMemcpy(pad_b, host, np * sizeof(double), hipMemcpyHostToDevice));
    unpack(s, pad_b, src, d, zero_ghosts);
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());   // the result is the last step: report a launch the runtime rejected
    HIP_TRY(hipMemcpy(host, pad_b, np * sizeof(double), hipMemcpyDeviceToHost));
    return PEPSMG_OK;
  }
  int scalar(int q, double* out) {
    HIP_TRY(hipMemcpyAsync(out, sc + q, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return PEPSMG_OK;
  }
};
using OneLevel = OneLevelT<double>;

// A device array of doubles freed when it goes out of scope, so every exit
// path of an entry point releases it.
struct DevBuf {
  double* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) hipFree(p);
  }
  int alloc(size_t n) {
    hipError_t e = hipMalloc(&p, n * sizeof(double));
    if (e != hipSuccess) {
      p = nullptr;
      return fail(PEPSMG_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(double), hipGetErrorString(e));
    }
    return PEPSMG_OK;
  }
};

// rbgs_solver on one level (src/rbgs_poisson_matrix.f90:73-171): kernel
// loop with the norm reduced on the device and tested on the host.
static int rbgs_solver_level(OneLevel& o, int maxit, double tol, double omega, int* iters, double* rel) {
  Ops<double> op{o.s};
  op.resid_norm(o.d, o.d.x, o.d.b, nullptr, o.partial, o.sc + 0);
  double r0, rn = 0;
  int rc = o.scalar(0, &r0);
  if (rc) return rc;
  rn = r0;
  int it;
  for (it = 0; it < maxit; it++) {
    op.rbgs(o.d, o.d.x, o.d.b, 1, omega, false);
    op.resid_norm(o.d, o.d.x, o.d.b, nullptr, o.partial, o.sc + 1);
    rc = o.scalar(1, &rn);
    if (rc) return rc;
    if (std::sqrt(rn / r0) <= tol) break;
  }
  if (iters) *iters = it;
  if (rel) *rel = std::sqrt(rn / r0);
  return PEPSMG_OK;
}

}  // namespace pepsmg

extern "C" {

int pepsmg_rbgs_iterator(const pepsmg_grid* grid, double* sol, const double* coeff, const double* rhs,
                         int32_t maxiteration, double omega) {
  OneLevel o;
  int rc = o.init(grid);
  if (!rc) rc = check_matrix(o.h, coeff);
  if (!rc) rc = check_ghosts_host(sol, o.d.nx, o.d.ny, o.d.nz, "sol");
  if (!rc) rc = o.upload(o.d.x, sol);
  if (!rc) rc = o.upload(o.d.b, rhs);
  if (rc) return rc;
  Ops<double> op{o.s};
  for (int it = 0; it < maxiteration; it++) op.fused_iter(o.d, omega, 0);
  return o.download(sol, o.d.x, 0);
}

int pepsmg_rbgs_iterator_fp32(const pepsmg_grid* grid, double* sol, const double* rhs, int32_t maxiteration,
                              double omega) {
  OneLevelT<float> o;
  int rc = o.init(grid);
  if (!rc) rc = check_ghosts_host(sol, o.d.nx, o.d.ny, o.d.nz, "sol");
  if (!rc) rc = o.upload(o.d.x, sol);
  if (!rc) rc = o.upload(o.d.b, rhs);
  if (rc) return rc;
  Ops<float> op{o.s};
  for (int it = 0; it < maxiteration; it++) op.fused_iter(o.d, (float)omega, 0);
  return o.download(sol, o.d.x, 0);
}

int pepsmg_rbgs_solver(const pepsmg_grid* grid, double* sol, const double* coeff, const double* rhs,
                       int32_t maxiteration, double tolerance, double omega, int32_t* iters, double* rel) {
  OneLevel o;
  int rc = o.init(grid);
  if (!rc) rc = check_matrix(o.h, coeff);
  if (!rc) rc = check_ghosts_host(sol, o.d.nx, o.d.ny, o.d.nz, "sol");
  if (!rc) rc = o.upload(o.d.x, sol);
  if (!rc) rc = o.upload(o.d.b, rhs);
  if (rc) return rc;
  int it = 0;
  double rl = 0;
  rc = rbgs_solver_level(o, maxiteration, tolerance, omega, &it, &rl);
  if (rc) return rc;
  if (iters) *iters = it;
  if (rel) *rel = rl;
  return o.download(sol, o.d.x, 0);
}

int pepsmg_cg_iterator(const pepsmg_grid* grid, double* sol, const double* coeff, const double* rhs,
                       int32_t maxiteration) {
  OneLevel o;
  int rc = o.init(grid);
  if (!rc) rc = check_matrix(o.h, coeff);
  if (!rc) rc = check_ghosts_host(sol, o.d.nx, o.d.ny, o.d.nz, "sol");
  if (!rc) rc = o.upload(o.d.x, sol);
  if (!rc) rc = o.upload(o.d.b, rhs);
  if (rc) return rc;
  Ops<double> op{o.s};
  DevBuf p, p2, ap;
  if ((rc = p.alloc(o.d.ncell)) || (rc = p2.alloc(o.d.ncell)) || (rc = ap.alloc(o.d.ncell))) return rc;
  cg_iterate(op, o.d, o.d.x, o.d.b, o.d.r, p.p, p2.p, ap.p, maxiteration, false, o.partial, o.sc);
  return o.download(sol, o.d.x, 0);
}

int pepsmg_cg_solver(const pepsmg_grid* grid, double* sol, const double* coeff, const double* rhs,
                     int32_t maxiteration, double tolerance, int32_t* iters, double* rel) {
  OneLevel o;
  int rc = o.init(grid);
  if (!rc) rc = check_matrix(o.h, coeff);
  if (!rc) rc = check_ghosts_host(sol, o.d.nx, o.d.ny, o.d.nz, "sol");
  if (!rc) rc = o.upload(o.d.x, sol);
  if (!rc) rc = o.upload(o.d.b, rhs);
  if (rc) return rc;
  Ops<double> op{o.s};
  DevBuf pb, apb;
  if ((rc = pb.alloc(o.d.ncell)) || (rc = apb.alloc(o.d.ncell))) return rc;
  double* p = pb.p;
  double* ap = apb.p;
  double* sc = o.sc;
  const long nrow = (long)o.d.ny * o.d.nz;
  const int eb = (int)std::max(1L, std::min<long>((nrow + CG_WAVES - 1) / CG_WAVES, RED_BLOCKS));
  const dim3 eblk(64 * CG_WAVES);
  // src/cg_poisson_matrix.f90:81-137.  (r, r) at the top of an iteration
  // (:106) is the (r, r) the previous iteration ended with (:123), the same
  // vector: sc[it & 1] holds it, sc[2] holds (Ap, p).
  op.resid_norm(o.d, o.d.x, o.d.b, o.d.r, o.partial, sc + 0);      // r = b - A x, rsd0tol (:94)
  HIP_TRY(hipMemcpyAsync(p, o.d.r, o.d.ncell * sizeof(double), hipMemcpyDeviceToDevice, o.s));
  double r0 = 0, t2 = 0;
  rc = o.scalar(0, &r0);
  int it;
  for (it = 0; it < maxiteration && !rc; it++) {
    const int cur = it & 1, nxt = (it + 1) & 1;
    op.apply(o.d, ap, p);                                          // :108
    op.dot(o.d, ap, p, o.partial, sc + 2);                         // :109
    hipLaunchKernelGGL((k_cg_xr<double>), dim3(eb), eblk, 0, o.s, o.d.x, o.d.r, p, ap, sc, cur, o.d.L(), false,
                       o.partial);                                 // :111-118
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, o.s, o.partial, eb, sc + nxt);   // :123
    rc = o.scalar(nxt, &t2);
    if (rc) break;
    if (std::sqrt(t2 / r0) < tolerance) break;                     // :126
    hipLaunchKernelGGL((k_cg_p<double>), dim3(eb), eblk, 0, o.s, p, o.d.r, sc, nxt, cur, o.d.L());
  }
  if (rc) return rc;
  if (iters) *iters = it;
  if (rel) *rel = std::sqrt(t2 / r0);
  return o.download(sol, o.d.x, 0);
}

int pepsmg_mv_mul(const pepsmg_grid* grid, double* y, const double* coeff, const double* x) {
  OneLevel o;
  int rc = o.init(grid);
  if (!rc) rc = check_matrix(o.h, coeff);
  if (!rc) rc = check_ghosts_host(x, o.d.nx, o.d.ny, o.d.nz, "x");
  if (!rc) rc = o.upload(o.d.x, x);
  if (rc) return rc;
  Ops<double> op{o.s};
  op.apply(o.d, o.d.r, o.d.x);
  return o.download(y, o.d.r, 1);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Scratch of pepsmg_vv_dot_device_noalloc: the per-block partial sums of one
// dot product (RED_BLOCKS doubles, 8 KiB per GPU, held until the process
// ends).  Every block writes its slot before the final reduction reads it,
// so nothing needs clearing between calls; calls that overlap on one GPU
// share it and race.  Host code never uses the symbol's name as an address:
// noalloc_scratch() returns the device address hipGetSymbolAddress gives for
// the current device.
namespace pepsmg {

constexpr int NOALLOC_SCRATCH_DOUBLES = RED_BLOCKS;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;

// Device address of scratch on the current device, looked up once per device
// (with that device current) and kept.
static int noalloc_scratch(double** out) {
  static std::mutex mu;
  static double* addr[NOALLOC_MAX_DEVICES] = {};
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(PEPSMG_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    addr[dev] = (double*)p;
  }
  *out = addr[dev];
  return PEPSMG_OK;
}

}  // namespace pepsmg

extern "C" {

int pepsmg_vv_dot_device_noalloc(int32_t nx, int32_t ny, int32_t nz, const double* d_x, const double* d_y,
                                 double* d_result, void* stream) {
  if (!d_x || !d_y || !d_result) return fail(PEPSMG_ERR_ARG, "NULL argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(PEPSMG_ERR_ARG, "grid %d x %d x %d is empty", nx, ny, nz);
  // rows (j, k) are numbered in int by the reduction kernel
  if ((long)ny * nz > 0x7fffffffL || nx > 0x7ffffffd)
    return fail(PEPSMG_ERR_ARG, "grid %d x %d x %d exceeds the 2^31-1 rows of cells the reduction numbers", nx, ny, nz);
  Lvl<double> L{};
  L.nx = nx;
  L.ny = ny;
  L.nz = nz;
  L.px = (long)nx + 2;
  L.pxy = L.px * ((long)ny + 2);
  const long o = 1 + L.px + L.pxy;   // cell (1,1,1) of the padded array
  // the block count of pepsmg_vv_dot (Ops::red_blocks), so both sum in the same order
  const long cells = (long)nx * ny * nz;
  const int nb = (int)std::max(1L, std::min<long>((cells + RED_THREADS - 1) / RED_THREADS, RED_BLOCKS));
  if (nb > NOALLOC_SCRATCH_DOUBLES) return fail(PEPSMG_ERR_ARG, "%d partial sums exceed the scratch", nb);
  double* part = nullptr;
  int rc = noalloc_scratch(&part);
  if (rc) return rc;
  (void)hipGetLastError();   // an error the caller's earlier HIP calls left is not this call's
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL((k_dot_partial<double, 0, false>), dim3(nb), dim3(RED_THREADS), 0, s, d_x + o, d_y + o, L, part,
                     (double*)nullptr);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, s, part, nb, d_result);
  HIP_TRY(hipGetLastError());   // a launch the runtime rejected
  return PEPSMG_OK;
}

int pepsmg_vv_dot(int32_t nx, int32_t ny, int32_t nz, const double* x, const double* y, double* result) {
  if (nx < 1 || ny < 1 || nz < 1 || !x || !y || !result) return fail(PEPSMG_ERR_ARG, "invalid argument");
  std::vector<double> ones[3];
  std::vector<double> dm[3], dg[3];
  int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; a++) {
    dm[a].assign(n[a] + 2, 1.0);
    dg[a].assign(n[a] + 2, 1.0);
  }
  pepsmg_grid g{nx, ny, nz, dm[0].data(), dg[0].data(), dm[1].data(), dg[1].data(), dm[2].data(), dg[2].data()};
  OneLevel o;
  int rc = o.init(&g);
  if (!rc) rc = o.upload(o.d.x, x);
  if (!rc) rc = o.upload(o.d.b, y);
  if (rc) return rc;
  Ops<double> op{o.s};
  op.dot(o.d, o.d.x, o.d.b, o.partial, o.sc);
  rc = o.scalar(0, result);
  if (rc) return rc;
  HIP_TRY(hipGetLastError());   // a launch the runtime rejected leaves its error here
  return PEPSMG_OK;
}

}  // extern "C"
