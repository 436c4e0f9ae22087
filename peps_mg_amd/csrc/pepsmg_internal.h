// pepsmg_internal.h -- host-side data structures shared by the library's
// translation units (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pepsmg.h"
#include "pepsmg_kernels.h"

#define PEPSMG_MAX_LEVELS 24

namespace pepsmg {

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
long pitch_for(int nx);

#define HIP_TRY(expr)                               \
  do {                                              \
    hipError_t e_ = (expr);                         \
    if (e_ != hipSuccess) return hip_fail(e_, #expr); \
  } while (0)

// Raises kernel kern's dynamic LDS limit to `bytes` on the current device.
// The attribute is per device, so it is set once per (kernel, device, size);
// a failure is returned and also stays pending for hipGetLastError, which
// the solve entry points check after their work has completed.
inline hipError_t set_dynamic_lds(const void* kern, size_t bytes) {
  static std::mutex mu;
  static std::set<std::tuple<const void*, int, size_t>> done;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_tuple(kern, dev, bytes);
  if (done.count(key)) return hipSuccess;
  e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) done.insert(key);
  return e;
}

// Host copy of one level's metrics and transfer weights (index 0..n+1).
struct HostLevel {
  int n[3] = {0, 0, 0};
  std::vector<double> dm[3], dg[3];     // dxm, dxg per axis
  std::vector<double> cw[3], ce[3];     // coefficients to the i-1 / i+1 neighbour
  std::vector<double> rw[3];            // restriction weights to the next level
  std::vector<double> pp[3], pn[3];     // prolongation weights from the next level
  int rs[3] = {1, 1, 1}, ps[3] = {1, 1, 1};
  int setup_coeffs();
};

template <typename T>
struct DevLevel {
  int nx = 0, ny = 0, nz = 0;
  long px = 0, pxy = 0;
  size_t ncell = 0;                       // allocated elements (pxy * nz)
  T *cw[3] = {nullptr, nullptr, nullptr}, *ce[3] = {nullptr, nullptr, nullptr};
  T *rw[3] = {nullptr, nullptr, nullptr}, *pp[3] = {nullptr, nullptr, nullptr}, *pn[3] = {nullptr, nullptr, nullptr};
  int rs[3] = {1, 1, 1}, ps[3] = {1, 1, 1};
  T *x = nullptr, *b = nullptr, *r = nullptr, *x2 = nullptr;
  bool fused_up = false;                  // the last fused smoother launch on this level marched up

  Lvl<T> L() const {
    Lvl<T> l;
    l.nx = nx; l.ny = ny; l.nz = nz; l.px = px; l.pxy = pxy;
    l.cwx = cw[0]; l.cex = ce[0]; l.cwy = cw[1]; l.cey = ce[1]; l.cwz = cw[2]; l.cez = ce[2];
    return l;
  }
  RestrictW<T> RW() const {
    RestrictW<T> w;
    w.wx = rw[0]; w.wy = rw[1]; w.wz = rw[2];
    w.sx = rs[0]; w.sy = rs[1]; w.sz = rs[2];
    return w;
  }
  ProlongW<T> PW() const {
    ProlongW<T> w;
    w.px_ = pp[0]; w.nx_ = pn[0]; w.py_ = pp[1]; w.ny_ = pn[1]; w.pz_ = pp[2]; w.nz_ = pn[2];
    w.sx = ps[0]; w.sy = ps[1]; w.sz = ps[2];
    return w;
  }
  long cells() const { return (long)nx * ny * nz; }
};

template <typename T>
struct Hier {
  std::vector<DevLevel<T>> lv;
  int build(const std::vector<HostLevel>& hl);
  void release();
};

int check_grid(const pepsmg_grid* g);
int build_hierarchy(const pepsmg_grid* g, int nlevel, int aggr, int cga_level, std::vector<HostLevel>& lv,
                    int lvc[3]);

// Extra per-level arrays of the CG smoother (p and A p).
template <typename T>
struct CgBufs {
  std::vector<T*> p, p2, ap;   // direction, the next direction, A p
  int ensure(Hier<T>& h) {
    if (!p.empty()) return 0;
    for (auto& d : h.lv) {
      T *a = nullptr, *b = nullptr, *c = nullptr;
      hipError_t e = hipMalloc(&a, d.ncell * sizeof(T));
      if (e == hipSuccess) e = hipMalloc(&b, d.ncell * sizeof(T));
      if (e == hipSuccess) e = hipMalloc(&c, d.ncell * sizeof(T));
      if (e != hipSuccess) return hip_fail(e, "hipMalloc (CG buffers)");
      hipMemset(a, 0, d.ncell * sizeof(T));
      hipMemset(b, 0, d.ncell * sizeof(T));
      hipMemset(c, 0, d.ncell * sizeof(T));
      p.push_back(a);
      ap.push_back(b);
      p2.push_back(c);
    }
    return 0;
  }
  void release() {
    for (auto q : p) hipFree(q);
    for (auto q : p2) hipFree(q);
    for (auto q : ap) hipFree(q);
    p.clear();
    p2.clear();
    ap.clear();
  }
};

// Measurement hook: brackets the plain (mode 0) fused smoother launches on
// one level, the finest level of the solve in progress, with HIP events on
// the solve stream; the durations are summed once the solve has finished.
// Launches beyond the event pool of one solve are not timed.
struct KernelTimer {
  bool on = false;
  const void* level = nullptr;   // the DevLevel whose plain fused launches are timed
  std::vector<hipEvent_t> ev;    // pairs (before, after)
  int used = 0;                  // events recorded in the solve in progress
  double total_ms = 0.0;
  long launches = 0;
  char name[96] = {0};           // the timed kernel's symbol
  bool wants(const void* lvl, int mode) const {
    return on && lvl == level && mode == 0 && used + 2 <= (int)ev.size();
  }
};

}  // namespace pepsmg

struct pepsmg_ctx {
  int nlevel = 0, ncycle = 0;
  int lvc[3] = {0, 0, 0};
  std::vector<pepsmg::HostLevel> host;
  pepsmg::Hier<double> h64;
  pepsmg::Hier<float> h32;
  pepsmg::CgBufs<double> cg64;
  pepsmg::CgBufs<float> cg32;
  pepsmg_options opt;
  hipStream_t stream = nullptr;
  size_t lds_max = 65536;        // dynamic LDS available to one workgroup
  // scratch
  double* d_partial = nullptr;   // reduction partials
  double* d_scalars = nullptr;   // reduction results / CG scalars
  int* d_flag = nullptr;         // ghost flag
  int* d_citers = nullptr;       // coarse iteration counts per cycle
  int citers_cap = 0;
  double* d_pad = nullptr;       // padded staging (host-pointer entry points)
  size_t pad_elems = 0;
  double* d_pad2 = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // convergence read-back: pinned host copy of d_scalars[0..7] and the event
  // after its copy, so the host can wait for the norm while later work runs
  double* h_scalars = nullptr;
  hipEvent_t ev_norm = nullptr;
  hipEvent_t ev_entry = nullptr;   // orders the own stream after the caller's default-stream work
  bool sol_out_done = false;       // the caller's sol holds the solution (stored there by the last step, or never left it)
  pepsmg::KernelTimer timer;       // measurement hook (pepsmg_kernel_timing)
  double* d_cinv = nullptr;        // coarsest level's inverse operator (mixed precision), built on first use
  // host copy of the ghost-cell flag of the solve in progress (read back with its first norm)
  int* ghost_flag_host() { return reinterpret_cast<int*>(h_scalars + 7); }

  int alloc_scratch();
  void release();
};
