// pepsmg_ops.h -- host-side launch helpers shared by the single-grid and
// batched solvers: block decompositions of the z-march kernels and one
// launcher per operation on one level.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>

#include "pepsmg_coarse.h"
#include "pepsmg_fused.h"
#include "pepsmg_internal.h"
#include "pepsmg_march.h"

namespace pepsmg {

// tile height of the fine-level residual kernels (norm, refinement update,
// CG steps)
constexpr int RESID_TY = 16;
// the refinement kernels' tile height on levels of 512 rows or more: 32
// rows halve the share of the one-row y ring re-read per tile (512^3 mixed
// solve 18.75 against 19.1 ms; two 256-thread blocks per CU by LDS either
// way).  Smaller levels keep 16: at about 2048 blocks, 32-row tiles leave
// z-chunks of 4 planes at 256^3, half of them halo.
constexpr int RESID_TY_BIG = 32;

// ---------------------------------------------------------------------------
// Launch helpers
static inline dim3 grid3(int nx, int ny, int nz, int bx, int by) {
  return dim3((unsigned)((nx + bx - 1) / bx), (unsigned)((ny + by - 1) / by), (unsigned)nz);
}

template <typename T>
struct Ops {
  hipStream_t s;
  int tile_y = 0;

  // x tile width of the fused smoother: 64 cells in FP64, 128 in FP32 (same
  // bytes per row)
  static constexpr int FTX = sizeof(T) == 8 ? 64 : 128;

  // Blocks of `kern` (launched with `threads` threads) that are resident on
  // the device at once: blocks per CU from the occupancy calculator times
  // the CU count (cached per kernel and device).
  static int resident(const void* kern, int threads) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_pair(kern, dev);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, 0) != hipSuccess) per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    return cache[key] = per * cus;
  }

  // Block decomposition of the fused smoother: TX x TY tiles marching
  // through z-chunks (even, so every chunk starts on an odd plane).  When
  // the tiles divide the resident block count, one chunk count fills the
  // device with exactly one wave of blocks (512^3: 0.58 against 0.63 ms in
  // FP64; 256^3: 73 against 100 us), chunks of at least 4 planes.
  // Otherwise about 2048 blocks, a multiple of 8 chunks (one z-slab per
  // XCD), chunks of at least 12 planes on levels of 192 or more planes and
  // of 2 below: levels of 128 planes and fewer are bound by the latency of
  // the march, not by bytes (FP32 smoother at 32^3 / 64^3: 6.8 / 7.5 us
  // against 8.2 / 8.7 us with chunks of 4 planes).
  // batched: the launch covers many problems at once, so it is bound by
  // bytes rather than latency and keeps chunks of at least 4 planes (12
  // from 96 planes on), which re-read fewer halo planes.
  static FusedGrid fused_grid(const DevLevel<T>& d, int TX, int TY, int resident_blocks = 0, bool batched = false) {
    FusedGrid G;
    G.gx = (d.nx + TX - 1) / TX;
    G.gy = (d.ny + TY - 1) / TY;
    long nxy = (long)G.gx * G.gy;
    long nch;
    if (resident_blocks > 0 && resident_blocks % nxy == 0 && d.nz / (resident_blocks / nxy) >= 4) {
      nch = resident_blocks / nxy;
    } else {
      const int min_planes = batched ? 4 : 2, long_from = batched ? 96 : 192;
      nch = std::max(1L, std::min<long>((2048 + nxy - 1) / nxy, (d.nz + min_planes - 1) / min_planes));
      if (nch < 8 && d.nz >= 64) nch = 8;   // one z-slab per XCD: tile rings hit the XCD's own L2
      if (nch >= 8) nch = (nch + 7) / 8 * 8;
      if (d.nz >= long_from) nch = std::min<long>(nch, std::max<long>(8, d.nz / 12 / 8 * 8));
    }
    int kc = (int)((d.nz + nch - 1) / nch);
    if (kc & 1) kc++;                      // even chunks: every chunk starts on an odd plane
    G.kchunk = kc;
    G.nch = (d.nz + G.kchunk - 1) / G.kchunk;
    // XCD-aware order for tiles of 2048 cells or more (512^3, one wave:
    // FP64 64 x 32 0.551 against 0.572 ms, FP32 128 x 32 0.289 against
    // 0.298 ms); 64 x 16 FP64 tiles are faster in chunk-fastest order
    // (0.595 against 0.620 ms)
    G.order = TX * TY >= 2048 ? 1 : 0;
    return G;
  }

  // resident blocks of the tile shape's plain kernel (every mode of a shape
  // uses the same decomposition, so mode 2's partial sums have one slot per
  // block of it)
  template <int TX, int TY, int PF>
  static int fused_resident() {
    return resident((const void*)k_rb_fused<T, TX, TY, 0, PF>, fused_threads(TX, TY));
  }

  // output buffer of the next fused launch (nullptr: d.x2)
  T* fused_out = nullptr;
  // modes 3 and 4 (FP64): the caller's padded arrays (pointers at cell
  // (1,1,1)) and their layout
  const T* caller_x = nullptr;
  const T* caller_b = nullptr;
  T* caller_sol = nullptr;
  Lvl<T> caller_lvl{};

  // DOWN: mode 0 marching each chunk from the top (k_rb_fused_down), same
  // decomposition and values
  template <int TX, int TY, int MODE, int PF, bool DOWN = false>
  void fused_launch(DevLevel<T>& d, T omega, double* partial) {
    FusedGrid G = fused_grid(d, TX, TY, fused_resident<TX, TY, PF>());
    unsigned nb = (unsigned)(G.gx * G.gy * G.nch);
    const T* xi = MODE == 3 ? caller_x : d.x;
    const T* b = MODE == 3 ? caller_b : d.b;
    T* ext = MODE == 3 ? d.b : (MODE == 4 ? caller_sol : nullptr);
    if constexpr (DOWN) {
      static_assert(MODE == 0 && PF == 1, "downward march: plain iterations, one plane of look-ahead");
      hipLaunchKernelGGL((k_rb_fused_down<T, TX, TY, MODE, PF>), dim3(nb), dim3(fused_threads(TX, TY)), 0, s,
                         fused_out ? fused_out : d.x2, xi, b, d.L(), omega, G, partial, d.L(), ext);
    } else {
      hipLaunchKernelGGL((k_rb_fused<T, TX, TY, MODE, PF>), dim3(nb), dim3(fused_threads(TX, TY)), 0, s,
                         fused_out ? fused_out : d.x2, xi, b, d.L(), omega, G, partial,
                         (MODE == 3 || MODE == 4) ? caller_lvl : d.L(), ext);
    }
  }

  template <int TX, int TY, int PF>
  void fused_mode(DevLevel<T>& d, T omega, int mode, double* partial, bool down) {
    if (mode == 1) fused_launch<TX, TY, 1, PF>(d, omega, nullptr);
    else if constexpr (sizeof(T) == 8) {
      if (mode == 2) fused_launch<TX, TY, 2, PF>(d, omega, partial);
      else if (mode == 3) fused_launch<TX, TY, 3, PF>(d, omega, partial);
      else if (mode == 4) fused_launch<TX, TY, 4, PF>(d, omega, nullptr);
      else fused_plain<TX, TY, PF>(d, omega, down);
    } else {
      fused_plain<TX, TY, PF>(d, omega, down);
    }
  }
  // The decomposition is sized by the upward kernel's residency, so a tile
  // shape whose downward kernel holds fewer blocks per CU (FP32 64 x 16: 137
  // against 127 registers) marches up.
  template <int TX, int TY, int PF>
  void fused_plain(DevLevel<T>& d, T omega, bool down) {
    if constexpr (PF == 1) {
      if (down && resident((const void*)k_rb_fused_down<T, TX, TY, 0, 1>, fused_threads(TX, TY)) ==
                      fused_resident<TX, TY, 1>()) {
        fused_launch<TX, TY, 0, 1, true>(d, omega, nullptr);
        return;
      }
    }
    d.fused_up = true;
    fused_launch<TX, TY, 0, PF>(d, omega, nullptr);
  }

  template <int TX>
  void fused_tx(DevLevel<T>& d, T omega, int mode, int ty, double* partial, bool down) {
    if (ty == 32) {
      fused_mode<TX, 32, 1>(d, omega, mode, partial, down);
    } else if (ty == 16) {
      if (prefetch == 2) fused_mode<TX, 16, 2>(d, omega, mode, partial, false);
      else fused_mode<TX, 16, 1>(d, omega, mode, partial, down);
    } else {
      if (prefetch == 2) fused_mode<TX, 8, 2>(d, omega, mode, partial, false);
      else fused_mode<TX, 8, 1>(d, omega, mode, partial, down);
    }
  }
  template <int TX>
  int fused_resident_tx(int ty) const {
    if (ty == 32) return fused_resident<TX, 32, 1>();
    if (ty == 16) return prefetch == 2 ? fused_resident<TX, 16, 2>() : fused_resident<TX, 16, 1>();
    return prefetch == 2 ? fused_resident<TX, 8, 2>() : fused_resident<TX, 8, 1>();
  }
  // blocks of one fused launch on level d (partial sums of mode 2)
  int fused_blocks(const DevLevel<T>& d) const {
    int tx = tx_used(d), ty = ty_used(d);
    int res = (sizeof(T) == 4 && tx == 64) ? fused_resident_tx<64>(ty) : fused_resident_tx<FTX>(ty);
    FusedGrid G = fused_grid(d, tx, ty, res);
    return G.gx * G.gy * G.nch;
  }

  // tile sizes in use: width 64 cells in FP64 and 128 in FP32 (the same
  // bytes per row) unless set
  int tile_x = 0;
  // Unless set, the tile follows the level: on levels of 256 rows or more
  // 64 x 32 (FP64) and 128 x 32 (FP32), below 64 x 8, whose smaller blocks
  // keep more CUs busy (FP32 solves: 256^3 4.6 against 5.2 ms, 128^3 2.2
  // against 3.0 ms; FP64 128^3 4.0 against 4.3 ms).  FP64 512^3 solve with
  // 64 x 32 tiles in XCD-aware order: 45.4 against 47.5 ms with 64 x 16.
  // FP32 levels of 128 rows whose width is a multiple of 128 (whole tiles)
  // take 128 x 32 too (smoother at 128^3: 13.3 against 17.5 us; at 96^3 and
  // 192^3, with partial tiles, 64 x 8 stays faster: 10.5 against 12.4 us
  // and 26.7 against 31.4 us).
  static bool big_level(const DevLevel<T>& d) {
    return d.ny >= 256 || (sizeof(T) == 4 && d.ny >= 128 && d.nx % 128 == 0);
  }
  static int default_tx(const DevLevel<T>& d) { return (sizeof(T) == 4 && big_level(d)) ? FTX : 64; }
  static int default_ty(const DevLevel<T>& d) { return big_level(d) ? 32 : 8; }
  int tx_used(const DevLevel<T>& d) const {
    if (sizeof(T) == 8) return 64;
    return tile_x ? (tile_x == 64 ? 64 : FTX) : default_tx(d);
  }
  int ty_used(const DevLevel<T>& d) const { return tile_y ? tile_y : default_ty(d); }

  // One fused red+black iteration x -> x2, then swap.  mode 0 plain, 1 zero
  // start (x taken as 0), 2 (FP64) plain with the partial sums of
  // (b - A x)^2 of the x it starts from written to partial[0 .. fused_blocks);
  // FP64 on the caller's arrays (set_caller): 3 as 2 with x and b read from
  // them and b stored into d.b, 4 as 0 with x also stored into the caller's sol.
  // prefetch: planes of look-ahead of the global loads.
  int prefetch = 1;
  // the caller's padded (0:nx+1, 0:ny+1, 0:nz+1) arrays on level d, for modes 3 and 4
  void set_caller(const DevLevel<T>& d, T* sol, const T* rhs) {
    caller_lvl = d.L();
    caller_lvl.px = d.nx + 2;
    caller_lvl.pxy = (long)(d.nx + 2) * (d.ny + 2);
    const long o = 1 + caller_lvl.px + caller_lvl.pxy;
    caller_sol = sol + o;
    caller_x = sol + o;
    caller_b = rhs + o;
  }

  // measurement hook: plain launches on the timer's level are bracketed by events
  KernelTimer* timer = nullptr;
  void timer_begin(const DevLevel<T>& d, int mode, bool& timed) {
    timed = timer && timer->wants(&d, mode);
    if (timed) hipEventRecord(timer->ev[timer->used], s);
  }
  void timer_end(bool timed) {
    if (!timed) return;
    hipEventRecord(timer->ev[timer->used + 1], s);
    timer->used += 2;
  }
  // A plain iteration that follows an upward launch on the same level
  // marches down, so that it first reads the planes the previous launch
  // touched last; every other fused launch marches up.
  bool alternate = true;
  bool next_down(DevLevel<T>& d, int mode) {
    const bool down = alternate && mode == 0 && d.fused_up && (ty_used(d) == 32 || prefetch == 1);
    d.fused_up = !down;
    return down;
  }
  void fused_iter(DevLevel<T>& d, T omega, int mode, double* partial = nullptr) {
    int ty = ty_used(d);
    const bool down = next_down(d, mode);
    bool timed;
    timer_begin(d, mode, timed);
    if (sizeof(T) == 4 && tx_used(d) == 64) fused_tx<64>(d, omega, mode, ty, partial, down);
    else fused_tx<FTX>(d, omega, mode, ty, partial, down);
    timer_end(timed);
    std::swap(d.x, d.x2);
  }
  // one fused iteration (mode 0) from d.x into out; no swap
  void fused_iter_into(DevLevel<T>& d, T omega, T* out) {
    fused_out = out;
    int ty = ty_used(d);
    const bool down = next_down(d, 0);
    bool timed;
    timer_begin(d, 0, timed);
    if (sizeof(T) == 4 && tx_used(d) == 64) fused_tx<64>(d, omega, 0, ty, nullptr, down);
    else fused_tx<FTX>(d, omega, 0, ty, nullptr, down);
    timer_end(timed);
    fused_out = nullptr;
  }
  // symbol of the plain fused kernel fused_iter launches on level d
  void fused_name(const DevLevel<T>& d, char* buf, size_t n) const {
    const int ty = ty_used(d);
    std::snprintf(buf, n, "k_rb_fused<%s, %d, %d, 0, %d>", sizeof(T) == 8 ? "double" : "float", tx_used(d), ty,
                  ty == 32 ? 1 : prefetch);   // 32-row tiles load one plane ahead
  }

  void smooth_colour(DevLevel<T>& d, const T* b, T* x, T omega, int colour, bool zero) {
    int hx = (d.nx + 1) / 2;
    dim3 g = grid3(hx, d.ny, d.nz, 64, 4), blk(64, 4, 1);
    if (zero)
      hipLaunchKernelGGL((k_rb_colour<T, true>), g, blk, 0, s, x, b, d.L(), omega, colour);
    else
      hipLaunchKernelGGL((k_rb_colour<T, false>), g, blk, 0, s, x, b, d.L(), omega, colour);
  }

  // nu red/black iterations on (x, b); zero_start: x is taken as 0 on entry
  void rbgs(DevLevel<T>& d, T* x, const T* b, int nu, T omega, bool zero_start) {
    if (nu <= 0) {
      if (zero_start) hipMemsetAsync(x, 0, d.ncell * sizeof(T), s);
      return;
    }
    for (int it = 0; it < nu; it++) {
      smooth_colour(d, b, x, omega, 0, zero_start && it == 0);
      smooth_colour(d, b, x, omega, 1, false);
    }
  }

  void residual(DevLevel<T>& d, T* r, const T* x, const T* b) {
    hipLaunchKernelGGL((k_residual<T>), grid3(d.nx, d.ny, d.nz, 64, 4), dim3(64, 4, 1), 0, s, r, x, b, d.L());
  }

  void apply(DevLevel<T>& d, T* y, const T* x) {
    hipLaunchKernelGGL((k_apply<T>), grid3(d.nx, d.ny, d.nz, 64, 4), dim3(64, 4, 1), 0, s, y, x, d.L());
  }

  // tile width of the restriction and prolongation z-march kernels: 64 cells
  // in FP64; in FP32 128 (512-byte rows either way) on levels where that
  // still leaves z-chunks of 8 or more planes (512^3: 20.8 against 21.0 ms
  // for the mixed solve), 64 below (shorter chunks re-read more halo planes)
  static constexpr int MTX = sizeof(T) == 8 ? 64 : 128;

  // z-march decomposition: TX x TY tiles, about 2048 blocks, chunk count a
  // multiple of 8 when >= 8; even chunks for the restriction kernel.
  static MarchGrid march_grid(const DevLevel<T>& d, int TY, bool even, int TX = 64) {
    MarchGrid G;
    G.gx = (d.nx + TX - 1) / TX;
    G.gy = (d.ny + TY - 1) / TY;
    long nxy = (long)G.gx * G.gy;
    long nch = std::max(1L, std::min<long>((2048 + nxy - 1) / nxy, (d.nz + 3) / 4));
    if (nch < 8 && d.nz >= 64) nch = 8;
    if (nch >= 8) nch = (nch + 7) / 8 * 8;
    int kc = (int)((d.nz + nch - 1) / nch);
    if (even && (kc & 1)) kc++;
    G.kchunk = kc;
    G.nch = (d.nz + kc - 1) / kc;
    return G;
  }

  void residual_restrict(DevLevel<T>& f, DevLevel<T>& c, const T* x, const T* b) {
    if (f.rs[0] == 2 && f.rs[1] == 2 && f.rs[2] == 2) {
      MarchGrid G = march_grid(f, 16, true, MTX);
      if (MTX == 64 || G.kchunk < 8) {
        G = march_grid(f, 16, true, 64);
          hipLaunchKernelGGL((k_restrict_march<T, 64>), dim3((unsigned)(G.gx * G.gy * G.nch)),
                           dim3(march_threads<64>()), 0, s, c.b, x, b, f.L(), c.L(), f.RW(), G);
      } else {
        hipLaunchKernelGGL((k_restrict_march<T, MTX>), dim3((unsigned)(G.gx * G.gy * G.nch)),
                           dim3(march_threads<MTX>()), 0, s, c.b, x, b, f.L(), c.L(), f.RW(), G);
      }
      return;
    }
    hipLaunchKernelGGL((k_restrict<T, true>), grid3(c.nx, c.ny, c.nz, 64, 4), dim3(64, 4, 1), 0, s, c.b, x, b,
                       f.L(), c.L(), f.RW());
  }

  // |b - A x|^2 into out[0] (partial sums per block, then one block); r
  // stored when non-null.
  void resid_norm_march(DevLevel<T>& d, const T* x, const T* b, T* r, double* partial, double* out) {
    MarchGrid G = march_grid(d, RESID_TY, false);
    int nb = G.gx * G.gy * G.nch;
    if (nb > PARTIAL_CAP) {
      resid_norm(d, x, b, r, partial, out);
      return;
    }
    if (r)
      hipLaunchKernelGGL((k_resid_march<T, RESID_TY, 1>), dim3((unsigned)nb), dim3(MARCH_THREADS), 0, s, x,
                         (const float*)nullptr, b, r, (float*)nullptr, (T*)nullptr, d.L(), G, partial);
    else
      hipLaunchKernelGGL((k_resid_march<T, RESID_TY, 0>), dim3((unsigned)nb), dim3(MARCH_THREADS), 0, s, x,
                         (const float*)nullptr, b, (T*)nullptr, (float*)nullptr, (T*)nullptr, d.L(), G, partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, s, partial, nb, out);
  }

  // prolongation march: TX x 16 fine tiles, chunks of coarse planes
  static MarchGrid prolong_grid(const DevLevel<T>& f, const DevLevel<T>& c, int TX = 64) {
    MarchGrid G;
    G.gx = (f.nx + TX - 1) / TX;
    G.gy = (f.ny + 15) / 16;
    long nxy = (long)G.gx * G.gy;
    long nch = std::max(1L, std::min<long>((2048 + nxy - 1) / nxy, (c.nz + 1) / 2));
    if (nch >= 8) nch = (nch + 7) / 8 * 8;
    G.kchunk = (int)((c.nz + nch - 1) / nch);
    G.nch = (c.nz + G.kchunk - 1) / G.kchunk;
    return G;
  }

  void prolong_add(DevLevel<T>& f, DevLevel<T>& c, T* xf) {
    if (f.ps[0] == 2 && f.ps[1] == 2 && f.ps[2] == 2) {
      MarchGrid G = prolong_grid(f, c, MTX);
      if (MTX == 64 || 2 * G.kchunk < 8) {   // chunks count coarse planes
        G = prolong_grid(f, c, 64);
          hipLaunchKernelGGL((k_prolong_march<T, 64>), dim3((unsigned)(G.gx * G.gy * G.nch)),
                           dim3(march_threads<64>()), 0, s, xf, c.x, f.L(), c.L(), f.PW(), G);
      } else {
        hipLaunchKernelGGL((k_prolong_march<T, MTX>), dim3((unsigned)(G.gx * G.gy * G.nch)),
                           dim3(march_threads<MTX>()), 0, s, xf, c.x, f.L(), c.L(), f.PW(), G);
      }
      return;
    }
    hipLaunchKernelGGL((k_prolong_add<T>), grid3(f.nx, f.ny, f.nz, 64, 4), dim3(64, 4, 1), 0, s, xf, c.x, f.L(),
                       c.L(), f.PW());
  }

  int red_blocks(const DevLevel<T>& d) {
    long n = d.cells();
    long b = (n + RED_THREADS - 1) / RED_THREADS;
    return (int)std::max(1L, std::min<long>(b, RED_BLOCKS));
  }

  // out[0] = sum x*y over the interior
  void dot(DevLevel<T>& d, const T* x, const T* y, double* partial, double* out) {
    int nb = red_blocks(d);
    hipLaunchKernelGGL((k_dot_partial<T, 0, false>), dim3(nb), dim3(RED_THREADS), 0, s, x, y, d.L(), partial,
                       (T*)nullptr);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, s, partial, nb, out);
  }

  // out[0] = |b - A x|^2; optionally r = b - A x
  void resid_norm(DevLevel<T>& d, const T* x, const T* b, T* r, double* partial, double* out) {
    int nb = red_blocks(d);
    if (r)
      hipLaunchKernelGGL((k_dot_partial<T, 1, true>), dim3(nb), dim3(RED_THREADS), 0, s, x, b, d.L(), partial, r);
    else
      hipLaunchKernelGGL((k_dot_partial<T, 1, false>), dim3(nb), dim3(RED_THREADS), 0, s, x, b, d.L(), partial,
                         (T*)nullptr);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, s, partial, nb, out);
  }
};


// Arguments of the persistent coarse kernel for levels s..L of hierarchy h:
// compact LDS layout of every level's x and b (16-byte aligned), the
// coarsest level in FP64.  Returns false if it does not fit in LDS.
template <typename T>
inline bool pc_build(int L, Hier<double>& h64, size_t lds_max, Hier<T>& h, int s, int nu, int nu_post, T omega,
                     double tol,
                     int* slot, PCArgs<T>& A, size_t& lds) {
  int nlev = L - s + 1;
  if (s < 1 || nlev > PC_MAXLEV) return false;
  std::memset(&A, 0, sizeof A);
  A.nlev = nlev;
  A.nu = nu;
  A.nu_post = nu_post;
  A.omega = omega;
  A.tol = tol;
  A.maxit = 1000;
  A.iters_slot = slot;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = (off + bytes + 15) / 16 * 16;
    return o;
  };
  for (int q = 0; q < nlev; q++) {
    DevLevel<T>& d = h.lv[s + q];
    Lvl<T> l = d.L();
    l.px = d.nx;
    l.pxy = (long)d.nx * d.ny;
    A.lv[q] = l;
    if (q < nlev - 1) {
      size_t n = (size_t)d.cells();
      A.xoff[q] = (int)(take(n * sizeof(T)) / sizeof(T));
      A.boff[q] = (int)(take(n * sizeof(T)) / sizeof(T));
      A.coff[q] = (int)(take((size_t)pc_stage_elems(d.nx, d.ny, d.nz, true) * sizeof(T)) / sizeof(T));
      A.rw[q] = d.RW();
      A.pw[q] = d.PW();
    }
  }
  if (nlev > 1) A.roff = (int)(take((size_t)h.lv[s].cells() * sizeof(T)) / sizeof(T));
  DevLevel<double>& cd = h64.lv[L];
  Lvl<double> lc = cd.L();
  lc.px = cd.nx;
  lc.pxy = (long)cd.nx * cd.ny;
  A.lc = lc;
  size_t nc = (size_t)cd.cells();
  A.xcoff = (int)(take(nc * 8) / 8);
  A.bcoff = (int)(take(nc * 8) / 8);
  A.xctoff = (int)(take(nc * sizeof(T)) / sizeof(T));
  A.ccoff = (int)(take((size_t)pc_stage_elems(cd.nx, cd.ny, cd.nz, false) * 8) / 8);
  A.gl = h.lv[s].L();
  A.gxs = h.lv[s].x;
  A.gbs = h.lv[s].b;
  lds = off;
  return off <= lds_max;
}

}  // namespace pepsmg
