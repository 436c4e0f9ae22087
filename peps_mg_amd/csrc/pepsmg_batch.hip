// pepsmg_batch.hip -- many independent Poisson solves in shared launches.
//
// A batch holds problems of different grid sizes, each with one or more
// right-hand sides; every (problem, right-hand side) pair is an item with
// its own level hierarchy.  Each operation of the V-cycle runs as ONE launch
// over all items: the launch's blocks are the concatenation of every item's
// blocks (item i owns blocks [block0_i, block0_{i+1})), and a block finds
// its item by binary search in a small device table.  Levels are aligned
// from the finest: down-leg step t smooths and restricts level t of every
// item that has a per-level level t; the persistent coarse kernel then runs
// one workgroup per item on levels s_i..L_i; the up-leg mirrors the down-leg.
// Items that have converged are switched off by a device flag that every
// batched kernel reads first, so one fixed launch sequence serves every step
// and is captured once into a hipGraph and replayed.
//
// Precision 0: FP64 V-cycles (the reference's multigrid_solve_vcycle,
// src/multigrid.f90:1907-1976), stop per item at rho < tol.  Precision 1:
// FP32 V-cycles inside FP64 iterative refinement (inner_cycles V-cycles per
// refinement step).  A replay is one V-cycle (FP64) or one refinement step
// (mixed); the host reads one counter of still-active items per replay.
// The values an item gets are exactly those of its single-grid solve with
// the same options (same kernels, same order), except that the final
// refinement step always runs inner_cycles cycles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <vector>

#include "../../include/pepsmg.h"
#include "pepsmg_internal.h"
#include "pepsmg_ops.h"

namespace pepsmg {

// ---------------------------------------------------------------------------
// Batched kernels: a table entry per item, sorted by block0.
template <typename Item>
__device__ __forceinline__ int batch_find(const Item* __restrict__ it, int n, int bid) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (it[mid].block0 <= bid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T>
struct FusedItem {
  int block0, item;
  T* xo;
  const T* xi;
  const T* b;
  Lvl<T> L;
  FusedGrid G;
};

template <typename T>
struct RestrictItem {
  int block0, item;
  T* bc;
  const T* x;
  const T* b;
  Lvl<T> F, C;
  RestrictW<T> W;
  MarchGrid G;
};

template <typename T>
struct ProlongItem {
  int block0, item;
  T* xf;
  const T* xc;
  Lvl<T> F, C;
  ProlongW<T> W;
  MarchGrid G;
};

// residual norm (KIND 0) or refinement update (KIND 2) on the fine level
struct ResidItem {
  int block0, item;
  const double* x;
  const float* e;
  const double* b;
  float* r32;
  double* xnew;
  Lvl<double> L;
  MarchGrid G;
  double* partial;
};

template <typename T, int TX, int TY, int MODE, int PF>
__global__ void __launch_bounds__(fused_threads(TX, TY))
kb_rb_fused(const FusedItem<T>* __restrict__ items, int n, const int* __restrict__ active, T omega) {
  const int q = batch_find(items, n, blockIdx.x);
  const FusedItem<T>& it = items[q];
  if (!active[it.item]) return;
  rb_fused_block<T, TX, TY, MODE, PF>(blockIdx.x - it.block0, it.xo, it.xi, it.b, it.L, omega, it.G, nullptr, it.L,
                                      (T*)nullptr);
}

template <typename T>
__global__ void __launch_bounds__(MARCH_THREADS)
kb_restrict(const RestrictItem<T>* __restrict__ items, int n, const int* __restrict__ active) {
  const int q = batch_find(items, n, blockIdx.x);
  const RestrictItem<T>& it = items[q];
  if (!active[it.item]) return;
  restrict_march_block<T, 64>(blockIdx.x - it.block0, it.bc, it.x, it.b, it.F, it.C, it.W, it.G);
}

template <typename T>
__global__ void __launch_bounds__(MARCH_THREADS)
kb_prolong(const ProlongItem<T>* __restrict__ items, int n, const int* __restrict__ active) {
  const int q = batch_find(items, n, blockIdx.x);
  const ProlongItem<T>& it = items[q];
  if (!active[it.item]) return;
  prolong_march_block<T, 64>(blockIdx.x - it.block0, it.xf, it.xc, it.F, it.C, it.W, it.G);
}

template <int KIND>
__global__ void __launch_bounds__(MARCH_THREADS)
kb_resid(const ResidItem* __restrict__ items, int n, const int* __restrict__ active) {
  const int q = batch_find(items, n, blockIdx.x);
  const ResidItem& it = items[q];
  if (!active[it.item]) return;
  resid_march_block<double, RESID_TY, KIND>(blockIdx.x - it.block0, it.x, it.e, it.b, (double*)nullptr, it.r32, it.xnew,
                                     it.L, it.G, it.partial, it.L, it.L, (double*)nullptr, it.L);
}

template <typename T>
__global__ void __launch_bounds__(PC_THREADS) kb_coarse(const PCArgs<T>* __restrict__ args, const int* __restrict__ active) {
  if (!active[blockIdx.x]) return;
  coarse_vcycle_block<T>(args[blockIdx.x]);
}

// out[i] = sum of item i's partial sums (same reduction order as k_dot_final)
__global__ void __launch_bounds__(RED_THREADS)
kb_dot_final(const double* __restrict__ partial, const int* __restrict__ poff, const int* __restrict__ pn,
             double* __restrict__ out, const int* __restrict__ active) {
  __shared__ double sh[RED_THREADS / 64];
  const int i = blockIdx.x;
  if (!active[i]) return;
  double acc = 0.0;
  for (int c = threadIdx.x; c < pn[i]; c += blockDim.x) acc += partial[poff[i] + c];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[i] = s;
}

// After a step: rho_i = sqrt(out_i / res0_i) into the item's history; items
// that reached rho < tol or their step cap, or whose rho is no longer
// finite, are switched off.  init: the first call records res0 (an item
// with res0 = 0 is already solved, one with a non-finite res0 has a NaN or
// infinity in its input; neither runs).
// nact receives the number of items still active.
__global__ void __launch_bounds__(256)
kb_check(const double* __restrict__ out, double* __restrict__ res0, double* __restrict__ hist, int hcap,
         int* __restrict__ active, int* __restrict__ steps, int* __restrict__ nact, int n, double tol, int maxsteps,
         int init) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (active[i]) {
      double v = out[i];
      if (init) {
        res0[i] = v;
        steps[i] = 0;
        if (!(v > 0.0)) active[i] = 0;
      } else {
        double rho = sqrt(v / res0[i]);
        int s = steps[i];
        if (s < hcap) hist[(long)i * hcap + s] = rho;
        steps[i] = s + 1;
        if (rho < tol || s + 1 >= maxsteps || !isfinite(rho)) active[i] = 0;
      }
    }
    if (active[i]) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) *nact = cnt;
}

__global__ void kb_activate(int* __restrict__ active, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) active[i] = 1;
}

// ---------------------------------------------------------------------------
// Host side
struct BItem {
  int prob = 0;
  int nlevel = 0;
  int s = 0;              // first level of the persistent coarse kernel
  int tile_x = 0, tile_y = 0;   // fused smoother tile of this item's problem (0: options)
  Hier<double> h64;
  Hier<float> h32;
  double* x_after_init = nullptr;   // fine FP64 x buffer holding the iterate after the initial update
};

// A recorded launch sequence with its item tables; tables live in one
// device arena uploaded before the sequence runs or is captured.
struct Sched {
  std::vector<char> host;
  std::vector<std::function<void(hipStream_t, const char*)>> ops;
  template <typename X>
  size_t put(const std::vector<X>& v) {
    size_t off = (host.size() + 255) / 256 * 256;
    host.resize(off + v.size() * sizeof(X));
    std::memcpy(host.data() + off, v.data(), v.size() * sizeof(X));
    return off;
  }
};

}  // namespace pepsmg

struct pepsmg_batch {
  int nprob = 0, nitems = 0, ncycle = 0;
  std::vector<std::vector<pepsmg::HostLevel>> host;
  std::vector<pepsmg::BItem> items;
  pepsmg_options opt;
  hipStream_t stream = nullptr, cap = nullptr;
  size_t lds_max = 65536;
  int* d_active = nullptr;
  int* d_steps = nullptr;
  int* d_nact = nullptr;
  int* d_flag = nullptr;
  int* d_poff = nullptr;
  int* d_pn = nullptr;
  double* d_out = nullptr;
  double* d_res0 = nullptr;
  double* d_hist = nullptr;
  double* d_partial = nullptr;
  int hcap = 0;
  // schedules: [0] and [1] differ only in which FP64 buffer the mixed
  // refinement reads / writes
  pepsmg::Sched sched[2];
  char* d_arena = nullptr;
  size_t arena_bytes = 0;
  size_t arena_off[2] = {0, 0};
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  bool built = false;
  int key_nu = -1;
  double key_tol = -1, key_omega = 0;
  pepsmg_options key_opt;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_entry = nullptr;   // orders the own stream after the caller's default-stream work

  void drop_schedule() {
    for (int p = 0; p < 2; p++) {
      if (exec[p]) hipGraphExecDestroy(exec[p]);
      if (graph[p]) hipGraphDestroy(graph[p]);
      exec[p] = nullptr;
      graph[p] = nullptr;
      sched[p] = pepsmg::Sched();
    }
    built = false;
  }
  void release() {
    drop_schedule();
    for (auto& it : items) {
      it.h64.release();
      it.h32.release();
    }
    items.clear();
    hipFree(d_active);
    hipFree(d_steps);
    hipFree(d_nact);
    hipFree(d_flag);
    hipFree(d_poff);
    hipFree(d_pn);
    hipFree(d_out);
    hipFree(d_res0);
    hipFree(d_hist);
    hipFree(d_partial);
    hipFree(d_arena);
    d_active = d_steps = d_nact = d_flag = d_poff = d_pn = nullptr;
    d_out = d_res0 = d_hist = d_partial = nullptr;
    d_arena = nullptr;
    if (ev0) hipEventDestroy(ev0);
    if (ev_entry) hipEventDestroy(ev_entry);
    ev_entry = nullptr;
    if (ev1) hipEventDestroy(ev1);
    ev0 = ev1 = nullptr;
    if (stream) hipStreamDestroy(stream);
    if (cap) hipStreamDestroy(cap);
    stream = cap = nullptr;
  }
};

namespace pepsmg {

// fine-level residual-march grid (same decomposition as the single-grid solver)
static MarchGrid resid_grid(const DevLevel<double>& f) { return Ops<double>::march_grid(f, RESID_TY, false); }

// Record nu fused iterations on level l of every item that has it (per-level
// levels only), swapping each item's x / x2 as the launches do.  Items are
// grouped by fused-smoother tile shape, one launch per group.  zero_first:
// the first iteration starts from x = 0.
template <typename T, int TX, int TY>
static void rec_smooth_group(pepsmg_batch* B, Sched& S, std::vector<Hier<T>*>& H, const std::vector<int>& members,
                             int l, bool zero, T omega) {
  std::vector<FusedItem<T>> tab;
  int nb = 0;
  for (int i : members) {
    DevLevel<T>& d = H[i]->lv[l];
    FusedItem<T> e;
    e.block0 = nb;
    e.item = i;
    e.xo = d.x2;
    e.xi = d.x;
    e.b = d.b;
    e.L = d.L();
    e.G = Ops<T>::fused_grid(d, TX, TY, 0, true);
    nb += e.G.gx * e.G.gy * e.G.nch;
    tab.push_back(e);
    std::swap(d.x, d.x2);
  }
  size_t off = S.put(tab);
  int n = (int)tab.size();
  int* act = B->d_active;
  S.ops.push_back([=](hipStream_t st, const char* dev) {
    const FusedItem<T>* p = (const FusedItem<T>*)(dev + off);
    const dim3 nt(fused_threads(TX, TY));
    if (zero) hipLaunchKernelGGL((kb_rb_fused<T, TX, TY, 1, 1>), dim3(nb), nt, 0, st, p, n, act, omega);
    else hipLaunchKernelGGL((kb_rb_fused<T, TX, TY, 0, 1>), dim3(nb), nt, 0, st, p, n, act, omega);
  });
}

template <typename T>
static void rec_smooth(pepsmg_batch* B, Sched& S, std::vector<Hier<T>*>& H, int l, int nu, T omega,
                       bool zero_first) {
  // tile shape per item: width 64 (FP64 always) or 128 (FP32), height 8, 16 or 32; unset, the
  // single-grid solver's default for a level of this size
  std::vector<int> g[2][3];
  for (int i = 0; i < B->nitems; i++) {
    if (l >= B->items[i].s) continue;
    int tx = B->items[i].tile_x ? B->items[i].tile_x : B->opt.tile_x;
    int ty = B->items[i].tile_y ? B->items[i].tile_y : B->opt.tile_y;
    if (!tx) tx = Ops<T>::default_tx(H[i]->lv[l]);
    if (!ty) ty = Ops<T>::default_ty(H[i]->lv[l]);
    int wx = (sizeof(T) == 4 && tx != 64) ? 1 : 0;
    g[wx][ty == 32 ? 2 : ty == 16 ? 1 : 0].push_back(i);
  }
  for (int it = 0; it < nu; it++) {
    bool z = zero_first && it == 0;
    if (!g[0][0].empty()) rec_smooth_group<T, 64, 8>(B, S, H, g[0][0], l, z, omega);
    if (!g[0][1].empty()) rec_smooth_group<T, 64, 16>(B, S, H, g[0][1], l, z, omega);
    if (!g[0][2].empty()) rec_smooth_group<T, 64, 32>(B, S, H, g[0][2], l, z, omega);
    if (sizeof(T) == 4) {
      if (!g[1][0].empty()) rec_smooth_group<T, Ops<T>::FTX, 8>(B, S, H, g[1][0], l, z, omega);
      if (!g[1][1].empty()) rec_smooth_group<T, Ops<T>::FTX, 16>(B, S, H, g[1][1], l, z, omega);
      if (!g[1][2].empty()) rec_smooth_group<T, Ops<T>::FTX, 32>(B, S, H, g[1][2], l, z, omega);
    }
  }
}

// residual of level l restricted into level l+1's b, for items with l < s
template <typename T>
static void rec_restrict(pepsmg_batch* B, Sched& S, std::vector<Hier<T>*>& H, int l) {
  std::vector<RestrictItem<T>> tab;
  int nb = 0;
  for (int i = 0; i < B->nitems; i++) {
    if (l >= B->items[i].s) continue;
    DevLevel<T>& f = H[i]->lv[l];
    DevLevel<T>& c = H[i]->lv[l + 1];
    RestrictItem<T> e;
    e.block0 = nb;
    e.item = i;
    e.bc = c.b;
    e.x = f.x;
    e.b = f.b;
    e.F = f.L();
    e.C = c.L();
    e.W = f.RW();
    e.G = Ops<T>::march_grid(f, 16, true);
    nb += e.G.gx * e.G.gy * e.G.nch;
    tab.push_back(e);
  }
  if (tab.empty()) return;
  size_t off = S.put(tab);
  int n = (int)tab.size();
  int* act = B->d_active;
  S.ops.push_back([=](hipStream_t st, const char* dev) {
    hipLaunchKernelGGL((kb_restrict<T>), dim3(nb), dim3(MARCH_THREADS), 0, st,
                       (const RestrictItem<T>*)(dev + off), n, act);
  });
}

// x_l += P x_{l+1}, for items with l < s
template <typename T>
static void rec_prolong(pepsmg_batch* B, Sched& S, std::vector<Hier<T>*>& H, int l) {
  std::vector<ProlongItem<T>> tab;
  int nb = 0;
  for (int i = 0; i < B->nitems; i++) {
    if (l >= B->items[i].s) continue;
    DevLevel<T>& f = H[i]->lv[l];
    DevLevel<T>& c = H[i]->lv[l + 1];
    ProlongItem<T> e;
    e.block0 = nb;
    e.item = i;
    e.xf = f.x;
    e.xc = c.x;
    e.F = f.L();
    e.C = c.L();
    e.W = f.PW();
    e.G = Ops<T>::prolong_grid(f, c);
    nb += e.G.gx * e.G.gy * e.G.nch;
    tab.push_back(e);
  }
  if (tab.empty()) return;
  size_t off = S.put(tab);
  int n = (int)tab.size();
  int* act = B->d_active;
  S.ops.push_back([=](hipStream_t st, const char* dev) {
    hipLaunchKernelGGL((kb_prolong<T>), dim3(nb), dim3(MARCH_THREADS), 0, st, (const ProlongItem<T>*)(dev + off), n,
                       act);
  });
}

template <typename T>
static void rec_coarse(pepsmg_batch* B, Sched& S, std::vector<Hier<T>*>& H, int nu, int nu_post, T omega,
                       double tol) {
  std::vector<PCArgs<T>> tab(B->nitems);
  size_t lds = 0;
  for (int i = 0; i < B->nitems; i++) {
    BItem& it = B->items[i];
    size_t li = 0;
    pc_build(it.nlevel, it.h64, B->lds_max, *H[i], it.s, nu, nu_post, omega, tol, (int*)nullptr, tab[i], li);
    lds = std::max(lds, li);
  }
  size_t off = S.put(tab);
  int n = B->nitems;
  int* act = B->d_active;
  S.ops.push_back([=](hipStream_t st, const char* dev) {
    hipLaunchKernelGGL((kb_coarse<T>), dim3(n), dim3(PC_THREADS), lds, st, (const PCArgs<T>*)(dev + off), act);
  });
}

// one V-cycle over every item of hierarchy set H (src/multigrid.f90:1931-1967)
template <typename T>
static void rec_vcycle(pepsmg_batch* B, Sched& S, std::vector<Hier<T>*>& H, int nu, int nu_post, T omega,
                       double tol, bool fine_zero) {
  int D = 0;
  for (auto& it : B->items) D = std::max(D, it.s);
  for (int l = 0; l < D; l++) {
    rec_smooth<T>(B, S, H, l, nu, omega, l > 0 || fine_zero);
    rec_restrict<T>(B, S, H, l);
  }
  rec_coarse<T>(B, S, H, nu, nu_post, omega, tol);
  for (int l = D - 1; l >= 0; l--) {
    rec_prolong<T>(B, S, H, l);
    rec_smooth<T>(B, S, H, l, nu_post, omega, false);
  }
}

// fine-level residual norm (KIND 0: of x) or refinement update (KIND 2:
// xnew = x + e, r32, norm), then per-item sums
template <int KIND>
static void rec_resid(pepsmg_batch* B, Sched& S, bool swap) {
  std::vector<ResidItem> tab;
  int nb = 0;
  long pofs = 0;
  for (int i = 0; i < B->nitems; i++) {
    BItem& it = B->items[i];
    DevLevel<double>& f = it.h64.lv[0];
    ResidItem e;
    e.block0 = nb;
    e.item = i;
    e.x = f.x;
    e.b = f.b;
    e.e = KIND == 2 ? it.h32.lv[0].x : nullptr;
    e.r32 = KIND == 2 ? it.h32.lv[0].b : nullptr;
    e.xnew = KIND == 2 ? f.x2 : nullptr;
    e.L = f.L();
    e.G = resid_grid(f);
    e.partial = B->d_partial + pofs;
    int nbi = e.G.gx * e.G.gy * e.G.nch;
    nb += nbi;
    pofs += nbi;
    tab.push_back(e);
    if (swap) std::swap(f.x, f.x2);
  }
  size_t off = S.put(tab);
  int n = (int)tab.size();
  int* act = B->d_active;
  double* partial = B->d_partial;
  int* poff = B->d_poff;
  int* pn = B->d_pn;
  double* out = B->d_out;
  S.ops.push_back([=](hipStream_t st, const char* dev) {
    hipLaunchKernelGGL((kb_resid<KIND>), dim3(nb), dim3(MARCH_THREADS), 0, st, (const ResidItem*)(dev + off), n, act);
    hipLaunchKernelGGL(kb_dot_final, dim3(n), dim3(RED_THREADS), 0, st, partial, poff, pn, out, act);
  });
}

static void rec_check(pepsmg_batch* B, Sched& S, double tol, int maxsteps, int init) {
  double *out = B->d_out, *res0 = B->d_res0, *hist = B->d_hist;
  int *act = B->d_active, *steps = B->d_steps, *nact = B->d_nact;
  int hcap = B->hcap, n = B->nitems;
  S.ops.push_back([=](hipStream_t st, const char*) {
    hipLaunchKernelGGL(kb_check, dim3(1), dim3(256), 0, st, out, res0, hist, hcap, act, steps, nact, n, tol, maxsteps,
                       init);
  });
}

}  // namespace pepsmg

using namespace pepsmg;

extern "C" {

int pepsmg_batch_create(int32_t nprob, const pepsmg_grid* grids, const int32_t* nlevel, const int32_t* nrhs,
                        int32_t ncycle, pepsmg_batch** out) {
  if (!out) return fail(PEPSMG_ERR_ARG, "NULL output pointer");
  *out = nullptr;
  if (nprob < 1 || !grids || !nlevel || !nrhs || ncycle < 1) return fail(PEPSMG_ERR_ARG, "invalid batch description");
  pepsmg_batch* B = new pepsmg_batch();
  B->nprob = nprob;
  B->ncycle = ncycle;
  pepsmg_default_options(&B->opt);
  B->host.resize(nprob);
  for (int p = 0; p < nprob; p++) {
    if (nrhs[p] < 1) {
      delete B;
      return fail(PEPSMG_ERR_ARG, "problem %d: nrhs < 1", p);
    }
    int lvc[3];
    int rc = build_hierarchy(&grids[p], nlevel[p], 0, 0, B->host[p], lvc);
    if (rc) {
      delete B;
      return rc;
    }
  }
  int dev = 0, lds = 0;
  hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 4096)
    B->lds_max = (size_t)lds - 2048;
  int n = 0;
  for (int p = 0; p < nprob; p++) n += nrhs[p];
  B->nitems = n;
  B->items.resize(n);
  int q = 0;
  for (int p = 0; p < nprob; p++)
    for (int r = 0; r < nrhs[p]; r++, q++) {
      BItem& it = B->items[q];
      it.prob = p;
      it.nlevel = nlevel[p];
      int rc = it.h64.build(B->host[p]);
      if (rc) {
        B->release();
        delete B;
        return rc;
      }
    }
  B->hcap = ncycle + 2;
  long partial = 0;
  std::vector<int> poff(n), pn(n);
  for (int i = 0; i < n; i++) {
    MarchGrid G = resid_grid(B->items[i].h64.lv[0]);
    poff[i] = (int)partial;
    pn[i] = G.gx * G.gy * G.nch;
    partial += pn[i];
  }
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess) {
      int rc = hip_fail(e, what);
      B->release();
      delete B;
      return rc;
    }
    return 0;
  };
  int rc;
  if ((rc = chk(hipMalloc(&B->d_active, n * sizeof(int)), "hipMalloc"))) return rc;
  if ((rc = chk(hipMalloc(&B->d_steps, n * sizeof(int)), "hipMalloc"))) return rc;
  if ((rc = chk(hipMalloc(&B->d_nact, sizeof(int)), "hipMalloc"))) return rc;
  if ((rc = chk(hipMalloc(&B->d_flag, sizeof(int)), "hipMalloc"))) return rc;
  if ((rc = chk(hipMalloc(&B->d_poff, n * sizeof(int)), "hipMalloc"))) return rc;
  if ((rc = chk(hipMalloc(&B->d_pn, n * sizeof(int)), "hipMalloc"))) return rc;
  if ((rc = chk(hipMalloc(&B->d_out, n * sizeof(double)), "hipMalloc"))) return rc;
  if ((rc = chk(hipMalloc(&B->d_res0, n * sizeof(double)), "hipMalloc"))) return rc;
  if ((rc = chk(hipMalloc(&B->d_hist, (size_t)n * B->hcap * sizeof(double)), "hipMalloc"))) return rc;
  if ((rc = chk(hipMalloc(&B->d_partial, std::max(1L, partial) * sizeof(double)), "hipMalloc"))) return rc;
  if ((rc = chk(hipMemcpy(B->d_poff, poff.data(), n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy"))) return rc;
  if ((rc = chk(hipMemcpy(B->d_pn, pn.data(), n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy"))) return rc;
  if ((rc = chk(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking), "hipStreamCreate"))) return rc;
  if ((rc = chk(hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking), "hipStreamCreate"))) return rc;
  if ((rc = chk(hipEventCreate(&B->ev0), "hipEventCreate"))) return rc;
  if ((rc = chk(hipEventCreate(&B->ev1), "hipEventCreate"))) return rc;
  if ((rc = chk(hipEventCreateWithFlags(&B->ev_entry, hipEventDisableTiming), "hipEventCreate"))) return rc;
  *out = B;
  return PEPSMG_OK;
}

int pepsmg_batch_set_options(pepsmg_batch* B, const pepsmg_options* opt) {
  if (!B || !opt) return fail(PEPSMG_ERR_ARG, "NULL argument");
  if (opt->smoother != 0) return fail(PEPSMG_ERR_ARG, "batched solves use the red/black smoother (smoother = 0)");
  if (opt->precision < 0 || opt->precision > 1) return fail(PEPSMG_ERR_ARG, "precision must be 0 or 1");
  if (opt->fused_smoother < 1 || opt->fused_smoother > 2)
    return fail(PEPSMG_ERR_ARG, "batched solves use the fused smoother (fused_smoother = 1 or 2)");
  if (opt->tile_y != 0 && opt->tile_y != 8 && opt->tile_y != 16 && opt->tile_y != 32)
    return fail(PEPSMG_ERR_ARG, "tile_y must be 0, 8, 16 or 32");
  if (opt->tile_x != 0 && opt->tile_x != 64 && opt->tile_x != 128) return fail(PEPSMG_ERR_ARG, "tile_x must be 0, 64 or 128");
  if (opt->inner_nu < 0 || opt->inner_nu_post < 0) return fail(PEPSMG_ERR_ARG, "inner_nu/inner_nu_post must be >= 0");
  if (!(opt->inner_omega >= 0.0 && opt->inner_omega < 2.0))
    return fail(PEPSMG_ERR_ARG, "inner_omega must be 0 (caller's omega) or in (0, 2)");
  if (opt->inner_cycles < 1 || opt->max_refine < 1) return fail(PEPSMG_ERR_ARG, "inner_cycles/max_refine must be >= 1");
  if (opt->coarse_switch < 1) return fail(PEPSMG_ERR_ARG, "batched solves need the persistent coarse kernel (coarse_switch > 0)");
  B->opt = *opt;
  B->drop_schedule();
  if (opt->precision == 1) {
    for (auto& it : B->items)
      if (it.h32.lv.empty()) {
        int rc = it.h32.build(B->host[it.prob]);
        if (rc) return rc;
      }
  }
  return PEPSMG_OK;
}

int32_t pepsmg_batch_size(const pepsmg_batch* B) { return B ? B->nitems : 0; }

int pepsmg_batch_set_tiles(pepsmg_batch* B, const int32_t* tile_x, const int32_t* tile_y) {
  if (!B) return fail(PEPSMG_ERR_ARG, "NULL argument");
  for (int p = 0; p < B->nprob; p++) {
    int tx = tile_x ? tile_x[p] : 0, ty = tile_y ? tile_y[p] : 0;
    if (tx != 0 && tx != 64 && tx != 128) return fail(PEPSMG_ERR_ARG, "problem %d: tile_x must be 0, 64 or 128", p);
    if (ty != 0 && ty != 8 && ty != 16 && ty != 32)
      return fail(PEPSMG_ERR_ARG, "problem %d: tile_y must be 0, 8, 16 or 32", p);
  }
  for (auto& it : B->items) {
    it.tile_x = tile_x ? tile_x[it.prob] : 0;
    it.tile_y = tile_y ? tile_y[it.prob] : 0;
  }
  B->drop_schedule();
  return PEPSMG_OK;
}

void pepsmg_batch_destroy(pepsmg_batch* B) {
  if (!B) return;
  B->release();
  delete B;
}

// switch level of every item for the given precision; false if some item's
// coarse levels do not fit one workgroup's LDS
static int batch_switch_levels(pepsmg_batch* B, int nu, double omega, double tol) {
  for (int i = 0; i < B->nitems; i++) {
    BItem& it = B->items[i];
    it.s = it.nlevel + 1;
    for (int l = 1; l <= it.nlevel; l++) {
      size_t lds = 0;
      bool ok;
      if (B->opt.precision == 1) {
        PCArgs<float> A;
        ok = it.h32.lv[l].cells() <= B->opt.coarse_switch &&
             pc_build(it.nlevel, it.h64, B->lds_max, it.h32, l, nu, nu, (float)omega, tol, (int*)nullptr, A, lds);
      } else {
        PCArgs<double> A;
        ok = it.h64.lv[l].cells() <= B->opt.coarse_switch &&
             pc_build(it.nlevel, it.h64, B->lds_max, it.h64, l, nu, nu, omega, tol, (int*)nullptr, A, lds);
      }
      if (ok) {
        it.s = l;
        break;
      }
    }
    if (it.s > it.nlevel)
      return fail(PEPSMG_ERR_ARG,
                  "item %d: no coarse level with at most coarse_switch cells fits one workgroup; use more levels", i);
    // the batched transfer kernels are the z-march ones: every level above
    // the switch level must halve all three axes
    for (int l = 0; l < it.s; l++) {
      const DevLevel<double>& d = it.h64.lv[l];
      if (d.rs[0] != 2 || d.rs[1] != 2 || d.rs[2] != 2)
        return fail(PEPSMG_ERR_ARG, "item %d: level %d does not halve every axis (batched solves need even sizes "
                                    "down to the coarse switch)", i, l);
    }
  }
  return PEPSMG_OK;
}

// Record both schedules, upload their tables, capture the graphs.
static int batch_build(pepsmg_batch* B, int nu, double tol, double omega) {
  B->drop_schedule();
  int rc = batch_switch_levels(B, nu, omega, tol);
  if (rc) return rc;
  const bool mixed = B->opt.precision == 1;
  const int maxsteps = mixed ? B->opt.max_refine : B->ncycle;
  if (mixed) {
    std::vector<Hier<float>*> H;
    for (auto& it : B->items) H.push_back(&it.h32);
    for (int p = 0; p < 2; p++) {
      Sched& S = B->sched[p];
      const int nui = B->opt.inner_nu > 0 ? B->opt.inner_nu : nu;
      const int nup = B->opt.inner_nu_post > 0 ? B->opt.inner_nu_post : nui;
      const float omi = (float)(B->opt.inner_omega > 0.0 ? B->opt.inner_omega : omega);
      for (int q = 0; q < B->opt.inner_cycles; q++) rec_vcycle<float>(B, S, H, nui, nup, omi, tol, q == 0);
      rec_resid<2>(B, S, true);
      rec_check(B, S, tol, maxsteps, 0);
    }
  } else {
    std::vector<Hier<double>*> H;
    for (auto& it : B->items) H.push_back(&it.h64);
    Sched& S = B->sched[0];
    rec_vcycle<double>(B, S, H, nu, nu, omega, tol, false);
    rec_resid<0>(B, S, false);
    rec_check(B, S, tol, maxsteps, 0);
  }
  // the recorders swapped pointers as the launches will; FP64 cycles swap an
  // even number of times, the mixed pair of steps too
  size_t total = 0;
  size_t offs[2];
  for (int p = 0; p < 2; p++) {
    offs[p] = (total + 255) / 256 * 256;
    total = offs[p] + B->sched[p].host.size();
  }
  if (total > B->arena_bytes) {
    hipFree(B->d_arena);
    B->d_arena = nullptr;
    HIP_TRY(hipMalloc(&B->d_arena, total));
    B->arena_bytes = total;
  }
  for (int p = 0; p < 2; p++)
    if (!B->sched[p].host.empty())
      HIP_TRY(hipMemcpy(B->d_arena + offs[p], B->sched[p].host.data(), B->sched[p].host.size(), hipMemcpyHostToDevice));
  HIP_TRY(set_dynamic_lds((const void*)kb_coarse<double>, B->lds_max));
  HIP_TRY(set_dynamic_lds((const void*)kb_coarse<float>, B->lds_max));
  for (int p = 0; p < (mixed ? 2 : 1); p++) {
    HIP_TRY(hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal));
    const char* base = B->d_arena + offs[p];
    for (auto& op : B->sched[p].ops) op(B->cap, base);
    HIP_TRY(hipStreamEndCapture(B->cap, &B->graph[p]));
    HIP_TRY(hipGraphInstantiate(&B->exec[p], B->graph[p], nullptr, nullptr, 0));
  }
  for (int p = 0; p < 2; p++) {
    B->arena_off[p] = offs[p];
    std::vector<char>().swap(B->sched[p].host);   // tables now live in the arena
  }
  B->key_nu = nu;
  B->key_tol = tol;
  B->key_omega = omega;
  B->key_opt = B->opt;
  B->built = true;
  return PEPSMG_OK;
}

static bool same_opt(const pepsmg_options& a, const pepsmg_options& b) { return std::memcmp(&a, &b, sizeof a) == 0; }

int pepsmg_batch_solve_device(pepsmg_batch* B, double* const* d_sol, const double* const* d_rhs, int32_t maxiteration,
                              double tolerance, double omega_sor, int32_t use_graph, void* stream,
                              pepsmg_stats* stats) {
  if (!B || !d_sol || !d_rhs) return fail(PEPSMG_ERR_ARG, "NULL argument");
  if (maxiteration < 1) return fail(PEPSMG_ERR_ARG, "batched solves need maxiteration >= 1");
  for (int i = 0; i < B->nitems; i++)
    if (!d_sol[i] || !d_rhs[i]) return fail(PEPSMG_ERR_ARG, "item %d: NULL array", i);
  hipStream_t st = stream ? (hipStream_t)stream : B->stream;
  if (!stream) {   // the own stream does not synchronize with the default stream: order it after the caller's work there
    HIP_TRY(hipEventRecord(B->ev_entry, (hipStream_t)0));
    HIP_TRY(hipStreamWaitEvent(st, B->ev_entry, 0));
  }
  const bool mixed = B->opt.precision == 1;
  const int n = B->nitems;
  // pointer state at the start of a solve: x in lv[0].x for every item
  if (!B->built || B->key_nu != maxiteration || B->key_tol != tolerance || B->key_omega != omega_sor ||
      !same_opt(B->key_opt, B->opt)) {
    int rc = batch_build(B, maxiteration, tolerance, omega_sor);
    if (rc) return rc;
  }
  (void)hipGetLastError();   // an error the caller's earlier HIP calls left is not this solve's
  HIP_TRY(hipEventRecord(B->ev0, st));
  // pack the caller's padded arrays; reject non-zero ghosts
  HIP_TRY(hipMemsetAsync(B->d_flag, 0, sizeof(int), st));
  for (int i = 0; i < n; i++) {
    DevLevel<double>& f = B->items[i].h64.lv[0];
    hipLaunchKernelGGL((k_pack<double>), dim3((f.nx + 2 + 63) / 64, (f.ny + 2 + 3) / 4, f.nz + 2), dim3(64, 4, 1), 0,
                       st, f.x, d_sol[i], f.L(), B->d_flag);
    hipLaunchKernelGGL((k_pack<double>), dim3((f.nx + 2 + 63) / 64, (f.ny + 2 + 3) / 4, f.nz + 2), dim3(64, 4, 1), 0,
                       st, f.b, d_rhs[i], f.L(), (int*)nullptr);
  }
  hipLaunchKernelGGL(kb_activate, dim3((n + 255) / 256), dim3(256), 0, st, B->d_active, n);
  int flag = 0;
  HIP_TRY(hipMemcpyAsync(&flag, B->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (flag) return fail(PEPSMG_ERR_GHOST, "a solution array has non-zero ghost cells (homogeneous Dirichlet required)");
  // initial residual (mixed: x2 = x + 0, r32 = b - A x, then x <-> x2)
  {
    Sched S;
    if (mixed) {
      for (auto& it : B->items)
        HIP_TRY(hipMemsetAsync(it.h32.lv[0].x, 0, it.h32.lv[0].ncell * sizeof(float), st));
      rec_resid<2>(B, S, true);
    } else {
      rec_resid<0>(B, S, false);
    }
    rec_check(B, S, tolerance, 1, 1);
    char* tab = nullptr;
    HIP_TRY(hipMalloc(&tab, std::max<size_t>(S.host.size(), 256)));
    HIP_TRY(hipMemcpyAsync(tab, S.host.data(), S.host.size(), hipMemcpyHostToDevice, st));
    for (auto& op : S.ops) op(st, tab);
    HIP_TRY(hipStreamSynchronize(st));
    hipFree(tab);
  }
  if (mixed)
    for (auto& it : B->items) it.x_after_init = it.h64.lv[0].x;
  // steps: replay until no item is active
  int nact = 0;
  HIP_TRY(hipMemcpy(&nact, B->d_nact, sizeof(int), hipMemcpyDeviceToHost));
  const int maxsteps = mixed ? B->opt.max_refine : B->ncycle;
  int step = 0;
  while (nact > 0 && step < maxsteps) {
    // schedule 1 was recorded from the state the initial update leaves
    // (x and x2 of the fine FP64 level exchanged), schedule 0 from the next
    int p = mixed ? ((step + 1) & 1) : 0;
    if (use_graph) {
      HIP_TRY(hipGraphLaunch(B->exec[p], st));
    } else {
      for (auto& op : B->sched[p].ops) op(st, B->d_arena + B->arena_off[p]);
    }
    HIP_TRY(hipMemcpyAsync(&nact, B->d_nact, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    step++;
  }
  // results: the fine x of each item; in mixed precision the buffer depends on
  // how many refinement updates the item took
  std::vector<int> steps(n);
  HIP_TRY(hipMemcpy(steps.data(), B->d_steps, n * sizeof(int), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++) {
    BItem& it = B->items[i];
    DevLevel<double>& f = it.h64.lv[0];
    double* xs = f.x;
    if (mixed) {
      double* other = (it.x_after_init == f.x) ? f.x2 : f.x;
      xs = (steps[i] & 1) ? other : it.x_after_init;
    }
    hipLaunchKernelGGL((k_unpack<double>), dim3((f.nx + 2 + 63) / 64, (f.ny + 2 + 3) / 4, f.nz + 2), dim3(64, 4, 1), 0,
                       st, d_sol[i], xs, f.L(), 0);
  }
  HIP_TRY(hipEventRecord(B->ev1, st));
  HIP_TRY(hipEventSynchronize(B->ev1));
  HIP_TRY(hipGetLastError());   // a launch the runtime rejected leaves its error here
  float ms = 0;
  hipEventElapsedTime(&ms, B->ev0, B->ev1);
  if (mixed) {
    // the initial update swapped every item's x / x2 once; swap back so the
    // recorded schedules (made from the pre-solve state) stay valid
    for (auto& it : B->items) std::swap(it.h64.lv[0].x, it.h64.lv[0].x2);
  }
  std::vector<double> res0(n), hist((size_t)n * B->hcap);
  HIP_TRY(hipMemcpy(res0.data(), B->d_res0, n * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(hist.data(), B->d_hist, hist.size() * sizeof(double), hipMemcpyDeviceToHost));
  int bad_input = -1, bad_rho = -1;
  for (int i = 0; i < n; i++) {
    int k = steps[i];
    if (!std::isfinite(res0[i]) && bad_input < 0) bad_input = i;
    if (k > 0 && !std::isfinite(hist[(size_t)i * B->hcap + std::min(k, B->hcap) - 1]) && bad_rho < 0) bad_rho = i;
  }
  if (stats) {
    for (int i = 0; i < n; i++) {
      pepsmg_stats& s = stats[i];
      int k = steps[i];
      s.res0 = res0[i];
      s.refine_steps = mixed ? k : 0;
      s.cycles = mixed ? k * B->opt.inner_cycles : k;
      s.rho = k > 0 ? hist[(size_t)i * B->hcap + std::min(k, B->hcap) - 1] : (res0[i] > 0 ? 1.0 : 0.0);
      s.converged = (k > 0 && s.rho < tolerance) || res0[i] == 0.0;
      if (s.history && s.history_cap > 0)
        for (int q = 0; q < std::min(k, std::min(s.history_cap, B->hcap)); q++)
          s.history[q] = hist[(size_t)i * B->hcap + q];
      s.device_ms = ms;
    }
  }
  if (bad_input >= 0)
    return fail(PEPSMG_ERR_ARG, "item %d: |b - A x0|^2 = %g is not finite (rhs or sol holds NaN or Inf)", bad_input,
                res0[bad_input]);
  if (bad_rho >= 0) return fail(PEPSMG_ERR_DIVERGED, "item %d: the relative residual is no longer finite", bad_rho);
  return PEPSMG_OK;
}

}  // extern "C"
