// pepsmg_kernels.h -- device kernels of the multigrid path (gfx950).
//
// Storage: every level keeps only its interior cells, x fastest, rows padded
// to a pitch of px elements (a multiple of 16) and planes of pxy = px*ny.
// Cell (i,j,k) (1-based, as in the reference) lives at
// (k-1)*pxy + (j-1)*px + (i-1).  The reference's ghost layer holds zeros
// (homogeneous Dirichlet) and is never stored: a neighbour outside
// [1,n] reads as 0.
//
// Arithmetic follows the reference expression by expression (same operand
// order, no fused multiply-add: the library is built with
// -ffp-contract=off), so in FP64 every smoother, residual and transfer value
// is bit-identical to the reference's.  The seven coefficients of
// src/matrix.f90:78-88 are rebuilt per cell from six 1-D arrays:
//   c1 = cwx[i], c2 = cex[i], c3 = cwy[j], c4 = cey[j], c5 = cwz[k], c6 = cez[k]
//   c0 = -(c1 + c2 + c3 + c4 + c5 + c6)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pepsmg {

template <typename T>
struct Lvl {
  int nx, ny, nz;
  long px, pxy;
  const T *cwx, *cex, *cwy, *cey, *cwz, *cez;
};

// Restriction weights from a fine level: per axis the reference's dxf array
// (src/multigrid.f90:1415-1458) and stride (1 or 2).
template <typename T>
struct RestrictW {
  const T *wx, *wy, *wz;
  int sx, sy, sz;
};

// Prolongation weights onto a fine level: per axis dxp/dxn
// (src/multigrid.f90:1594-1649) and stride.
template <typename T>
struct ProlongW {
  const T *px_, *nx_, *py_, *ny_, *pz_, *nz_;
  int sx, sy, sz;
};

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for
// every outstanding global load (its fence covers global memory), which
// would drain the register prefetch of the z-march kernels at each plane;
// their barriers only publish LDS writes.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Load of a read-only array element at a wave-uniform index (per-plane z
// coefficients and weights) through the constant address space, so it is
// a scalar load: it waits on the scalar/LDS counter, not on the vector
// memory counter that also tracks the plane prefetches in flight.  Only for
// arrays no kernel writes.
template <typename T>
__device__ __forceinline__ T ldc(const T* p, int i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Streaming accesses for data a kernel touches once (right-hand sides,
// outputs): marked non-temporal so they do not evict the x lines that
// neighbouring tiles re-read from L2.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void st_stream_pair(T* p, T a, T b) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  v2 o = {a, b};
  __builtin_nontemporal_store(o, reinterpret_cast<v2*>(p));
}

template <typename T>
__device__ __forceinline__ long lidx(const Lvl<T>& L, int i, int j, int k) {
  return (long)(k - 1) * L.pxy + (long)(j - 1) * L.px + (i - 1);
}

template <typename T>
__device__ __forceinline__ T ldz(const T* __restrict__ a, const Lvl<T>& L, int i, int j, int k) {
  return (i >= 1 && i <= L.nx && j >= 1 && j <= L.ny && k >= 1 && k <= L.nz) ? a[lidx(L, i, j, k)]
                                                                             : T(0);
}

// c0 in the reference's summation order (src/matrix.f90:82)
template <typename T>
__device__ __forceinline__ T diag(T c1, T c2, T c3, T c4, T c5, T c6) {
  return -(c1 + c2 + c3 + c4 + c5 + c6);
}

// A x at (i,j,k): mv_mul order, c0 term first (src/poisson_matrix_operator.f90:150-156)
template <typename T>
__device__ __forceinline__ T apply_at(const T* __restrict__ x, const Lvl<T>& L, int i, int j, int k) {
  T c1 = L.cwx[i], c2 = L.cex[i], c3 = L.cwy[j], c4 = L.cey[j], c5 = L.cwz[k], c6 = L.cez[k];
  T c0 = diag(c1, c2, c3, c4, c5, c6);
  return c0 * ldz(x, L, i, j, k) + c1 * ldz(x, L, i - 1, j, k) + c2 * ldz(x, L, i + 1, j, k) +
         c3 * ldz(x, L, i, j - 1, k) + c4 * ldz(x, L, i, j + 1, k) + c5 * ldz(x, L, i, j, k - 1) +
         c6 * ldz(x, L, i, j, k + 1);
}

// r = b - A x (src/multigrid.f90:1795-1805)
template <typename T>
__device__ __forceinline__ T resid_at(const T* __restrict__ x, const T* __restrict__ b, const Lvl<T>& L,
                                      int i, int j, int k) {
  T y = apply_at(x, L, i, j, k);
  return b[lidx(L, i, j, k)] - y;
}

// SOR update of one cell (src/rbgs_poisson_matrix.f90:244-250)
template <typename T>
__device__ __forceinline__ T sor_at(const T* __restrict__ x, const T* __restrict__ b, const Lvl<T>& L,
                                    int i, int j, int k, T omega, T xold) {
  T c1 = L.cwx[i], c2 = L.cex[i], c3 = L.cwy[j], c4 = L.cey[j], c5 = L.cwz[k], c6 = L.cez[k];
  T c0 = diag(c1, c2, c3, c4, c5, c6);
  T temp = c1 * ldz(x, L, i - 1, j, k) + c2 * ldz(x, L, i + 1, j, k) + c3 * ldz(x, L, i, j - 1, k) +
           c4 * ldz(x, L, i, j + 1, k) + c5 * ldz(x, L, i, j, k - 1) + c6 * ldz(x, L, i, j, k + 1);
  return omega * (b[lidx(L, i, j, k)] - temp) / c0 + (T(1) - omega) * xold;
}

// ---------------------------------------------------------------------------
// One colour of a red/black sweep, in place.  colour 0 = cells with i+j+k
// odd (1-based), the reference's first sweep (ista = 1 + mod(j+k,2),
// src/rbgs_poisson_matrix.f90:242); colour 1 the second (:260).
// ZERO: the sweep starts from x = 0 (the zeroed coarse-level x of
// src/multigrid.f90:1937); the colour is computed with zero neighbours and
// the other colour's cells are written as 0, so x need not be cleared first.
// Thread (tx, j, k) handles the tx-th cell of the colour in row (j,k).
template <typename T, bool ZERO>
__global__ void k_rb_colour(T* __restrict__ x, const T* __restrict__ b, Lvl<T> L, T omega, int colour) {
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > L.ny) return;
  int ista = 1 + ((j + k + colour) & 1);
  int i = ista + 2 * t;
  if (ZERO) {
    // other colour: plain zero store (cell ista-1+... of the pair)
    int io = (ista == 1) ? 2 + 2 * t : 1 + 2 * t;
    if (io <= L.nx) x[lidx(L, io, j, k)] = T(0);
  }
  if (i > L.nx) return;
  T c1 = L.cwx[i], c2 = L.cex[i], c3 = L.cwy[j], c4 = L.cey[j], c5 = L.cwz[k], c6 = L.cez[k];
  T c0 = diag(c1, c2, c3, c4, c5, c6);
  T v;
  if (ZERO) {
    T z = T(0);
    T temp = c1 * z + c2 * z + c3 * z + c4 * z + c5 * z + c6 * z;
    v = omega * (b[lidx(L, i, j, k)] - temp) / c0 + (T(1) - omega) * z;
  } else {
    v = sor_at(x, b, L, i, j, k, omega, x[lidx(L, i, j, k)]);
  }
  x[lidx(L, i, j, k)] = v;
}

// r = b - A x on the interior
template <typename T>
__global__ void k_residual(T* __restrict__ r, const T* __restrict__ x, const T* __restrict__ b, Lvl<T> L) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  if (i > L.nx || j > L.ny) return;
  r[lidx(L, i, j, k)] = resid_at(x, b, L, i, j, k);
}

// y = A x on the interior
template <typename T>
__global__ void k_apply(T* __restrict__ y, const T* __restrict__ x, Lvl<T> L) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  if (i > L.nx || j > L.ny) return;
  y[lidx(L, i, j, k)] = apply_at(x, L, i, j, k);
}

// ---------------------------------------------------------------------------
// Fused residual + restriction (src/multigrid.f90:1932-1933 / :1939-1940 with
// multigrid_restriction :1460-1494): each coarse cell averages the fine
// residual of its (up to) 8 children, weights vol_f = dxf*dyf*dzf.
template <typename T>
__device__ __forceinline__ void restrict_axis(const T* w, int s, int c, int& pz, int& pp, T& w1, T& w2) {
  pp = c * s;
  pz = pp - (s - 1);
  w1 = w[pz];   // slot 1 weight, applied to the value at pp
  w2 = w[pp];   // slot 2 weight, applied to the value at pz
}

template <typename T, bool FROM_RESIDUAL>
__global__ void k_restrict(T* __restrict__ bc, const T* __restrict__ xf, const T* __restrict__ bf,
                           Lvl<T> F, Lvl<T> Cc, RestrictW<T> W) {
  int ic = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int jc = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int kc = blockIdx.z + 1;
  if (ic > Cc.nx || jc > Cc.ny) return;
  int iz, ip, jz, jp, kz, kp;
  T ax1, ax2, ay1, ay2, az1, az2;
  restrict_axis(W.wx, W.sx, ic, iz, ip, ax1, ax2);
  restrict_axis(W.wy, W.sy, jc, jz, jp, ay1, ay2);
  restrict_axis(W.wz, W.sz, kc, kz, kp, az1, az2);
  T v111 = ax1 * ay1 * az1, v211 = ax2 * ay1 * az1, v121 = ax1 * ay2 * az1, v221 = ax2 * ay2 * az1;
  T v112 = ax1 * ay1 * az2, v212 = ax2 * ay1 * az2, v122 = ax1 * ay2 * az2, v222 = ax2 * ay2 * az2;
  T vol_c = v111 + v211 + v121 + v221 + v112 + v212 + v122 + v222;
  auto R = [&](int i, int j, int k) -> T {
    if (FROM_RESIDUAL) return resid_at(xf, bf, F, i, j, k);
    return bf[lidx(F, i, j, k)];
  };
  T s = v111 * R(ip, jp, kp) + v211 * R(iz, jp, kp) + v121 * R(ip, jz, kp) + v221 * R(iz, jz, kp) +
        v112 * R(ip, jp, kz) + v212 * R(iz, jp, kz) + v122 * R(ip, jz, kz) + v222 * R(iz, jz, kz);
  bc[lidx(Cc, ic, jc, kc)] = s / vol_c;
}

// ---------------------------------------------------------------------------
// Prolongation + correction (src/multigrid.f90:1958-1959, :1965-1966 with
// multigrid_prolongation_linear_on_nonuniform_grid :1676-1754): x_f += P x_c.
// Per axis a fine cell f takes its parent c with weight ws and one coarse
// neighbour nb with weight wn:
//   stride 2, f odd : ws = dxp(f),  wn = dxp(f+1), nb = c-1
//   stride 2, f even: ws = dxn(f),  wn = dxn(f-1), nb = c+1
//   stride 1        : ws = wn = dxn(f), nb = c   (the last of the reference's
//                     eight assignments to the cell wins)
// The eight terms are summed in the reference order (self/neighbour per axis:
// sss, nss, sns, ssn, nns, nsn, snn, nnn) and divided by the product of the
// per-axis weight sums.
template <typename T>
__device__ __forceinline__ void prolong_axis(const T* p, const T* n, int s, int f, int& c, int& nb, T& ws,
                                             T& wn) {
  if (s == 2) {
    c = (f + 1) >> 1;
    if (f & 1) {
      ws = p[f]; wn = p[f + 1]; nb = c - 1;
    } else {
      ws = n[f]; wn = n[f - 1]; nb = c + 1;
    }
  } else {
    c = f; ws = n[f]; wn = n[f]; nb = c;
  }
}

template <typename T>
__device__ __forceinline__ T prolong_at(const T* __restrict__ xc, const Lvl<T>& Cc, const ProlongW<T>& W,
                                        int i, int j, int k) {
  int ci, ni, cj, nj, ck, nk;
  T wsx, wnx, wsy, wny, wsz, wnz;
  prolong_axis(W.px_, W.nx_, W.sx, i, ci, ni, wsx, wnx);
  prolong_axis(W.py_, W.ny_, W.sy, j, cj, nj, wsy, wny);
  prolong_axis(W.pz_, W.nz_, W.sz, k, ck, nk, wsz, wnz);
  T num = wsx * wsy * wsz * ldz(xc, Cc, ci, cj, ck) + wnx * wsy * wsz * ldz(xc, Cc, ni, cj, ck) +
          wsx * wny * wsz * ldz(xc, Cc, ci, nj, ck) + wsx * wsy * wnz * ldz(xc, Cc, ci, cj, nk) +
          wnx * wny * wsz * ldz(xc, Cc, ni, nj, ck) + wnx * wsy * wnz * ldz(xc, Cc, ni, cj, nk) +
          wsx * wny * wnz * ldz(xc, Cc, ci, nj, nk) + wnx * wny * wnz * ldz(xc, Cc, ni, nj, nk);
  return num / ((wsx + wnx) * (wsy + wny) * (wsz + wnz));
}

template <typename T>
__global__ void k_prolong_add(T* __restrict__ xf, const T* __restrict__ xc, Lvl<T> F, Lvl<T> Cc,
                              ProlongW<T> W) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  if (i > F.nx || j > F.ny) return;
  long q = lidx(F, i, j, k);
  xf[q] = xf[q] + prolong_at(xc, Cc, W, i, j, k);
}

// ---------------------------------------------------------------------------
// Deterministic reductions: fixed grid, fixed per-thread order, fixed tree.
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;
constexpr int PARTIAL_CAP = 1 << 18;   // per-block partial sums of one reduction

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int q = 0; q < nw; q++) t += sh[q];
  }
  __syncthreads();
  return t;
}

// partial[block] = sum over the block's cells of x*y (MODE 0), or of r*r with
// r = b - A x (MODE 1, optionally storing r and/or a float copy of r)
template <typename T, int MODE, bool STORE_R>
__global__ void __launch_bounds__(RED_THREADS) k_dot_partial(const T* __restrict__ x, const T* __restrict__ y,
                                                             Lvl<T> L, double* __restrict__ partial,
                                                             T* __restrict__ rout) {
  __shared__ double sh[RED_THREADS / 64];
  // each wave walks whole rows (j,k); rows are dealt to waves round-robin
  const int lane = threadIdx.x & 63;
  const int nw = RED_THREADS / 64;
  const int nrow = L.ny * L.nz;
  double acc = 0.0;
  for (int row = blockIdx.x * nw + (threadIdx.x >> 6); row < nrow; row += gridDim.x * nw) {
    int j = row % L.ny + 1;
    int k = row / L.ny + 1;
    for (int i = lane + 1; i <= L.nx; i += 64) {
      if (MODE == 0) {
        long q = lidx(L, i, j, k);
        acc += (double)x[q] * (double)y[q];
      } else {
        T r = resid_at(x, y, L, i, j, k);
        if (STORE_R) rout[lidx(L, i, j, k)] = r;
        acc += (double)r * (double)r;
      }
    }
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

static __global__ void __launch_bounds__(RED_THREADS) k_dot_final(const double* __restrict__ partial, int n,
                                                           double* __restrict__ out) {
  __shared__ double sh[RED_THREADS / 64];
  double acc = 0.0;
  for (int c = threadIdx.x; c < n; c += blockDim.x) acc += partial[c];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) *out = s;
}

// ---------------------------------------------------------------------------
// Coarsest-level solve in one workgroup (src/multigrid.f90:1819-1851 ->
// rbgs_solver_poisson_matrix, src/rbgs_poisson_matrix.f90:47-179).  x is
// zeroed, then red/black iterations until sqrt(|r|^2/|r0|^2) <= tol or
// maxit.  A single cell is solved directly (x = b/c0).  For up to 64 cells
// the norms are summed by one thread in k,j,i order (the reference's
// single-thread order), so the iteration count and every value match the
// reference bit for bit.  iters_out receives the reference's loop variable.
constexpr int COARSE_THREADS = 256;

template <typename T>
__device__ double coarse_norm(const T* x, const T* b, const Lvl<T>& L, double* sh, double* seq) {
  int n = L.nx * L.ny * L.nz;
  if (n <= COARSE_THREADS && n <= 64) {
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      int i = c % L.nx + 1, j = (c / L.nx) % L.ny + 1, k = c / (L.nx * L.ny) + 1;
      double r = (double)resid_at(x, b, L, i, j, k);
      seq[c] = r * r;
    }
    __syncthreads();
    __shared__ double tot;
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int c = 0; c < n; c++) s = s + seq[c];
      tot = s;
    }
    __syncthreads();
    return tot;
  }
  double acc = 0.0;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    int i = c % L.nx + 1, j = (c / L.nx) % L.ny + 1, k = c / (L.nx * L.ny) + 1;
    double r = (double)resid_at(x, b, L, i, j, k);
    acc += r * r;
  }
  __shared__ double tot2;
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) tot2 = s;
  __syncthreads();
  return tot2;
}

template <typename T>
__device__ void coarse_colour(T* x, const T* b, const Lvl<T>& L, T omega, int colour) {
  int hx = (L.nx + 1) / 2;
  int n = hx * L.ny * L.nz;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    int t = c % hx, j = (c / hx) % L.ny + 1, k = c / (hx * L.ny) + 1;
    int i = 1 + ((j + k + colour) & 1) + 2 * t;
    if (i <= L.nx) {
      long q = lidx(L, i, j, k);
      x[q] = sor_at(x, b, L, i, j, k, omega, x[q]);
    }
  }
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(COARSE_THREADS) k_coarse_solve(T* __restrict__ x, const T* __restrict__ b,
                                                                 Lvl<T> L, T omega, double tol, int maxit,
                                                                 int* __restrict__ iters_out,
                                                                 int* __restrict__ iters_slot) {
  __shared__ double sh[COARSE_THREADS / 64];
  __shared__ double seq[64];
  int n = L.nx * L.ny * L.nz;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    int i = c % L.nx + 1, j = (c / L.nx) % L.ny + 1, k = c / (L.nx * L.ny) + 1;
    x[lidx(L, i, j, k)] = T(0);
  }
  __syncthreads();
  if (n == 1) {
    if (threadIdx.x == 0) {
      T c1 = L.cwx[1], c2 = L.cex[1], c3 = L.cwy[1], c4 = L.cey[1], c5 = L.cwz[1], c6 = L.cez[1];
      x[0] = b[0] / diag(c1, c2, c3, c4, c5, c6);
      if (iters_out) *iters_out = 0;
      if (iters_slot) *iters_slot = 0;
    }
    return;
  }
  double r0 = coarse_norm(x, b, L, sh, seq);
  int it;
  for (it = 0; it < maxit; it++) {
    coarse_colour(x, b, L, omega, 0);
    coarse_colour(x, b, L, omega, 1);
    double rn = coarse_norm(x, b, L, sh, seq);
    if (sqrt(rn / r0) <= tol) break;
  }
  if (threadIdx.x == 0) {
    if (iters_out) *iters_out = it;
    if (iters_slot) *iters_slot = it;
  }
}

// ---------------------------------------------------------------------------
// Layout conversion between the reference's padded arrays and the compact
// level storage.  pack also flags non-zero ghost cells.
template <typename T>
__global__ void k_pack(T* __restrict__ dst, const double* __restrict__ src, Lvl<T> L, int* __restrict__ ghost_flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;   // 0..nx+1
  int j = blockIdx.y * blockDim.y + threadIdx.y;   // 0..ny+1
  int k = blockIdx.z;                              // 0..nz+1
  if (i > L.nx + 1 || j > L.ny + 1) return;
  long s = (long)i + (long)(L.nx + 2) * ((long)j + (long)(L.ny + 2) * k);
  double v = src[s];
  bool ghost = (i == 0 || j == 0 || k == 0 || i == L.nx + 1 || j == L.ny + 1 || k == L.nz + 1);
  if (ghost) {
    if (ghost_flag && v != 0.0) atomicOr(ghost_flag, 1);
  } else {
    dst[lidx(L, i, j, k)] = (T)v;
  }
}

// flag = 1 if a padded (0:nx+1, 0:ny+1, 0:nz+1) array has a non-zero ghost
// cell.  Work items, visited by a grid-stride loop: the two z faces and the
// y-face rows of every interior plane, one cell each (consecutive items
// are consecutive addresses), then the two end cells of every interior row,
// one row each.
static __global__ void __launch_bounds__(256)
k_ghost_check(const double* __restrict__ a, int nx, int ny, int nz, int* __restrict__ flag) {
  const long px = nx + 2, pxy = px * (ny + 2);
  const long nface = 2 * pxy + 2L * nz * px;   // z faces, then rows j = 0 and ny+1 of planes 1..nz
  const long nrow = (long)ny * nz;              // interior rows
  bool bad = false;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nface + nrow; q += (long)gridDim.x * blockDim.x) {
    if (q < pxy) {
      bad |= a[q] != 0.0;                                       // plane k = 0
    } else if (q < 2 * pxy) {
      bad |= a[(long)(nz + 1) * pxy + (q - pxy)] != 0.0;        // plane k = nz + 1
    } else if (q < nface) {
      const long e = q - 2 * pxy, r = e / px, i = e - r * px;   // r: (plane, face row) pairs
      const long k = 1 + r / 2, j = (r & 1) ? ny + 1 : 0;
      bad |= a[k * pxy + j * px + i] != 0.0;
    } else {
      const long e = q - nface, k = 1 + e / ny, j = 1 + e % ny;
      const double* row = a + k * pxy + j * px;
      bad |= (row[0] != 0.0) | (row[nx + 1] != 0.0);
    }
  }
  if (bad) atomicOr(flag, 1);
}

template <typename T>
__global__ void k_unpack(double* __restrict__ dst, const T* __restrict__ src, Lvl<T> L, int zero_ghosts) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (i > L.nx + 1 || j > L.ny + 1) return;
  long s = (long)i + (long)(L.nx + 2) * ((long)j + (long)(L.ny + 2) * k);
  bool ghost = (i == 0 || j == 0 || k == 0 || i == L.nx + 1 || j == L.ny + 1 || k == L.nz + 1);
  if (ghost) {
    if (zero_ghosts) dst[s] = 0.0;
  } else {
    dst[s] = (double)src[lidx(L, i, j, k)];
  }
}

// ---------------------------------------------------------------------------
// Elementwise helpers (mixed precision and CG).
template <typename T>
__global__ void k_fill(T* __restrict__ a, long n, T v) {
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) a[c] = v;
}

}  // namespace pepsmg
