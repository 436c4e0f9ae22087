// pepsmg_coarse.h -- prolongation+correction z-march kernel and the
// persistent coarse-level kernel (gfx950).
#pragma once
#include "pepsmg_march.h"

namespace pepsmg {

// ---------------------------------------------------------------------------
// x_f += P x_c for a level that halves every axis
// (src/multigrid.f90:1958-1959 / :1965-1966 with
// multigrid_prolongation_linear_on_nonuniform_grid :1676-1754).  A block
// owns a TX x 16 fine tile (a TX/2 x 8 coarse tile; TX = 64, or 128 for
// FP32, with 4 TX threads); coarse planes with a
// one-cell ring are staged in LDS (4 slots), fine x planes stream through
// registers with the next plane's loads in flight.  Per axis the fine cell f
// takes its parent c with weight ws and the neighbour nb with weight wn
// (f odd: dxp(f), dxp(f+1), c-1; f even: dxn(f), dxn(f-1), c+1); the eight
// terms are summed in the reference order.
template <typename T, int TX>
__device__ __forceinline__ void prolong_march_block(int bid, T* __restrict__ xf, const T* __restrict__ xc,
                                                    const Lvl<T>& F, const Lvl<T>& Cc, const ProlongW<T>& W,
                                                    const MarchGrid& G) {
  constexpr int TP = TX / 2, NT = march_threads<TX>();
  constexpr int CX = TP + 2, CY = 10;                // coarse tile TX/2 x 8 + ring 1
  constexpr int NCL = (CX * CY + NT - 1) / NT;
  constexpr int NP = 2;                              // fine pairs per thread
  using V2 = typename Vec2<T>::type;
  __shared__ T sc[4][CY][CX];

  const int t = threadIdx.x;
  const int bz = bid % G.nch;
  const int rxy = bid / G.nch;
  const int i0 = 1 + (rxy % G.gx) * TX;              // fine tile origin
  const int j0 = 1 + (rxy / G.gx) * 16;
  const int ic0 = (i0 + 1) / 2, jc0 = (j0 + 1) / 2;  // coarse tile origin
  const int kc0 = 1 + bz * G.kchunk;                 // coarse planes of this chunk
  const int kc1 = min(kc0 + G.kchunk, Cc.nz + 1);

  // fine pairs: rows ly = t/TP + 8q, pair lp = t%TP -> cells i0+2lp (odd), i0+2lp+1
  int ly[NP];
  bool act[NP], two[NP];
  T wsa[NP], wna[NP], wsb[NP], wnb[NP], wsy[NP], wny[NP];
  int cjr[NP], njr[NP];   // coarse row (LDS) of parent / neighbour in y
  const int lp = t % TP;
  const int ia = i0 + 2 * lp;
  // x: cell ia (odd): parent ic, neighbour ic-1; cell ia+1 (even): parent ic, neighbour ic+1
  const int ic = (ia + 1) / 2;
  const int lcx = ic - ic0 + 1;                      // LDS column of the parent
#pragma unroll
  for (int q = 0; q < NP; q++) {
    ly[q] = t / TP + 8 * q;
    int gj = j0 + ly[q];
    act[q] = gj <= F.ny && ia <= F.nx;
    two[q] = act[q] && ia + 1 <= F.nx;
    wsa[q] = act[q] ? W.px_[ia] : T(0);
    wna[q] = act[q] ? W.px_[ia + 1] : T(0);
    wsb[q] = two[q] ? W.nx_[ia + 1] : T(0);
    wnb[q] = two[q] ? W.nx_[ia] : T(0);
    int jc = (gj + 1) / 2;
    bool jodd = gj & 1;
    wsy[q] = act[q] ? (jodd ? W.py_[gj] : W.ny_[gj]) : T(0);
    wny[q] = act[q] ? (jodd ? W.py_[gj + 1] : W.ny_[gj - 1]) : T(0);
    cjr[q] = jc - jc0 + 1;
    njr[q] = (jodd ? jc - 1 : jc + 1) - jc0 + 1;
  }

  // Coarse and fine planes load from clamped in-bounds addresses with no
  // branch around the loads; coarse values outside the domain are zeroed
  // when staged, fine values of cells outside are never stored.
  T cpre[NCL];
  int cq = 0;
  auto issue_c = [&](int kc) {
    cq = kc;
    const int kcc = min(max(kc, 1), Cc.nz);
#pragma unroll
    for (int q = 0; q < NCL; q++) {
      int e = min(t + q * NT, CX * CY - 1);
      int cy = e / CX, cx = e - cy * CX;
      int gi = min(max(ic0 - 1 + cx, 1), Cc.nx), gj = min(max(jc0 - 1 + cy, 1), Cc.ny);
      cpre[q] = xc[lidx(Cc, gi, gj, kcc)];
    }
  };
  auto stage_c = [&](int kc) {
    const bool vz = cq >= 1 && cq <= Cc.nz;
#pragma unroll
    for (int q = 0; q < NCL; q++) {
      int e = t + q * NT;
      if (e < CX * CY) {
        int cy = e / CX, cx = e - cy * CX;
        int gi = ic0 - 1 + cx, gj = jc0 - 1 + cy;
        bool v = vz && gi >= 1 && gi <= Cc.nx && gj >= 1 && gj <= Cc.ny;
        (&sc[kc & 3][0][0])[e] = v ? cpre[q] : T(0);
      }
    }
  };
  // fine planes are loaded two planes ahead: xpre[h] holds the next fine
  // plane of parity h (h = 0: odd, 1: even)
  V2 xpre[2][NP];
  auto issue_f = [&](V2* dst, int f) {
    const int fc = min(max(f, 1), F.nz);
#pragma unroll
    for (int q = 0; q < NP; q++) {
      // an odd column <= nx keeps the 2-element load inside the padded row
      int gi = act[q] ? ia : 1, gj = act[q] ? j0 + ly[q] : 1;
      dst[q] = ld_stream_v2(xf + lidx(F, gi, gj, fc));
    }
  };

  issue_c(kc0 - 1);
  stage_c(kc0 - 1);
  issue_c(kc0);
  stage_c(kc0);
  issue_c(kc0 + 1);
  stage_c(kc0 + 1);
  issue_c(kc0 + 2);
  issue_f(xpre[0], 2 * kc0 - 1);
  issue_f(xpre[1], 2 * kc0);
  // coarse-plane step, run twice per iteration so the compiler's wait
  // counts across the back edge stay exact
  auto step = [&](int kc) {
    lds_barrier();
#pragma unroll
    for (int h = 0; h < 2; h++) {
      int f = 2 * kc - 1 + h;                        // h = 0: odd fine plane, 1: even (F.nz = 2 Cc.nz)
      V2 xv[NP];
#pragma unroll
      for (int q = 0; q < NP; q++) xv[q] = xpre[h][q];
      issue_f(xpre[h], f + 2);
      T wsz = (h == 0) ? ldc(W.pz_, f) : ldc(W.nz_, f);
      T wnz = (h == 0) ? ldc(W.pz_, f + 1) : ldc(W.nz_, f - 1);
      T(*P)[CX] = sc[kc & 3];
      T(*N)[CX] = sc[(h == 0 ? kc - 1 : kc + 1) & 3];
#pragma unroll
      for (int q = 0; q < NP; q++) {
        if (!act[q]) continue;
        int cj = cjr[q], nj = njr[q];
        T sy = wsy[q], ny = wny[q];
        // first cell (odd i): parent column lcx, neighbour lcx-1
        {
          T sx = wsa[q], nx = wna[q];
          int ci = lcx, ni = lcx - 1;
          T num = sx * sy * wsz * P[cj][ci] + nx * sy * wsz * P[cj][ni] + sx * ny * wsz * P[nj][ci] +
                  sx * sy * wnz * N[cj][ci] + nx * ny * wsz * P[nj][ni] + nx * sy * wnz * N[cj][ni] +
                  sx * ny * wnz * N[nj][ci] + nx * ny * wnz * N[nj][ni];
          xv[q].x = xv[q].x + num / ((sx + nx) * (sy + ny) * (wsz + wnz));
        }
        if (two[q]) {   // second cell (even i): parent lcx, neighbour lcx+1
          T sx = wsb[q], nx = wnb[q];
          int ci = lcx, ni = lcx + 1;
          T num = sx * sy * wsz * P[cj][ci] + nx * sy * wsz * P[cj][ni] + sx * ny * wsz * P[nj][ci] +
                  sx * sy * wnz * N[cj][ci] + nx * ny * wsz * P[nj][ni] + nx * sy * wnz * N[cj][ni] +
                  sx * ny * wnz * N[nj][ci] + nx * ny * wnz * N[nj][ni];
          xv[q].y = xv[q].y + num / ((sx + nx) * (sy + ny) * (wsz + wnz));
        }
        long q0 = lidx(F, ia, j0 + ly[q], f);
        if (two[q]) st_stream_pair(xf + q0, xv[q].x, xv[q].y);
        else xf[q0] = xv[q].x;
      }
    }
    stage_c(kc + 2);
    issue_c(kc + 3);
  };
  int kc = kc0;
  for (; kc + 1 < kc1; kc += 2) {
    step(kc);
    step(kc + 1);
  }
  if (kc < kc1) step(kc);
}

template <typename T, int TX>
__global__ void __launch_bounds__(march_threads<TX>())
k_prolong_march(T* __restrict__ xf, const T* __restrict__ xc, Lvl<T> F, Lvl<T> Cc, ProlongW<T> W, MarchGrid G) {
  prolong_march_block<T, TX>(blockIdx.x, xf, xc, F, Cc, W, G);
}

// ---------------------------------------------------------------------------
// Persistent coarse-level kernel: one workgroup runs the part of the
// V-cycle below a switch level s entirely in LDS (src/multigrid.f90:1936-1962
// restricted to levels s..L): for each level down to L-1, nu red/black
// iterations from x = 0, residual and restriction; the coarsest solve
// (rbgs_solver, <= 1000 iterations, stop at sqrt(|r|^2/|r0|^2) <= tol, a
// single cell directly); then up again: prolongation + correction and nu
// iterations.  It reads b_s from global memory and writes x_s back.  The
// coarsest level is solved in FP64 whatever T is (T = float in the
// mixed-precision inner cycles).  Operations run in the reference order and
// the coarsest norm of up to 64 cells is summed sequentially, so FP64
// results are bit-identical to the per-level kernels and to the reference.
constexpr int PC_MAXLEV = 8;
constexpr int PC_THREADS = 1024;

template <typename T>
struct PCArgs {
  int nlev;                       // levels s..L handled: nlev = L - s + 1
  int nu;
  T omega;
  double tol;
  int maxit;
  Lvl<T> lv[PC_MAXLEV];           // compact LDS geometry (px = nx, pxy = nx*ny)
  Lvl<double> lc;                 // coarsest level, FP64 coefficients
  RestrictW<T> rw[PC_MAXLEV];     // level q -> q+1
  ProlongW<T> pw[PC_MAXLEV];      // level q <- q+1
  int xoff[PC_MAXLEV], boff[PC_MAXLEV];   // LDS offsets (elements of T)
  int xcoff, bcoff, xctoff;       // coarsest: FP64 x, b (in doubles), T copy of x (in T)
  Lvl<T> gl;                      // global geometry of level s
  T* gxs;                         // global x_s (output)
  const T* gbs;                   // global b_s (input)
  int* iters_slot;
};

__device__ __forceinline__ int pc_i(int c, int nx) { return c % nx + 1; }

template <typename T>
__device__ void pc_colour(T* x, const T* b, const Lvl<T>& L, T omega, int colour, bool zero) {
  int hx = (L.nx + 1) / 2;
  int n = hx * L.ny * L.nz;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    int tt = c % hx, j = (c / hx) % L.ny + 1, k = c / (hx * L.ny) + 1;
    int i = 1 + ((j + k + colour) & 1) + 2 * tt;
    if (i <= L.nx) {
      long q = lidx(L, i, j, k);
      if (zero) {
        T z = T(0);
        T c1 = L.cwx[i], c2 = L.cex[i], c3 = L.cwy[j], c4 = L.cey[j], c5 = L.cwz[k], c6 = L.cez[k];
        T temp = c1 * z + c2 * z + c3 * z + c4 * z + c5 * z + c6 * z;
        x[q] = omega * (b[q] - temp) / diag(c1, c2, c3, c4, c5, c6) + (T(1) - omega) * z;
      } else {
        x[q] = sor_at(x, b, L, i, j, k, omega, x[q]);
      }
    }
  }
  __syncthreads();
}

template <typename T>
__device__ void pc_zero(T* a, int n) {
  for (int c = threadIdx.x; c < n; c += blockDim.x) a[c] = T(0);
  __syncthreads();
}

template <typename T>
__device__ void pc_smooth(T* x, const T* b, const Lvl<T>& L, T omega, int nu, bool zero_start) {
  int n = L.nx * L.ny * L.nz;
  if (zero_start) pc_zero(x, n);   // zero-start colour 0 writes only its own colour
  for (int it = 0; it < nu; it++) {
    pc_colour(x, b, L, omega, 0, zero_start && it == 0);
    pc_colour(x, b, L, omega, 1, false);
  }
}

template <typename T, typename TC>
__device__ void pc_restrict(TC* bc, const T* x, const T* b, const Lvl<T>& F, int cnx, int cny, int cnz,
                            const RestrictW<T>& W) {
  int n = cnx * cny * cnz;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    int ic = c % cnx + 1, jc = (c / cnx) % cny + 1, kc = c / (cnx * cny) + 1;
    int iz, ip, jz, jp, kz, kp;
    T ax1, ax2, ay1, ay2, az1, az2;
    restrict_axis(W.wx, W.sx, ic, iz, ip, ax1, ax2);
    restrict_axis(W.wy, W.sy, jc, jz, jp, ay1, ay2);
    restrict_axis(W.wz, W.sz, kc, kz, kp, az1, az2);
    T v111 = ax1 * ay1 * az1, v211 = ax2 * ay1 * az1, v121 = ax1 * ay2 * az1, v221 = ax2 * ay2 * az1;
    T v112 = ax1 * ay1 * az2, v212 = ax2 * ay1 * az2, v122 = ax1 * ay2 * az2, v222 = ax2 * ay2 * az2;
    T vol_c = v111 + v211 + v121 + v221 + v112 + v212 + v122 + v222;
    T s = v111 * resid_at(x, b, F, ip, jp, kp) + v211 * resid_at(x, b, F, iz, jp, kp) +
          v121 * resid_at(x, b, F, ip, jz, kp) + v221 * resid_at(x, b, F, iz, jz, kp) +
          v112 * resid_at(x, b, F, ip, jp, kz) + v212 * resid_at(x, b, F, iz, jp, kz) +
          v122 * resid_at(x, b, F, ip, jz, kz) + v222 * resid_at(x, b, F, iz, jz, kz);
    bc[(ic - 1) + cnx * ((jc - 1) + cny * (kc - 1))] = (TC)(s / vol_c);
  }
  __syncthreads();
}

template <typename T>
__device__ void pc_prolong_add(T* xf, const T* xc, const Lvl<T>& F, const Lvl<T>& Cc, const ProlongW<T>& W) {
  int n = F.nx * F.ny * F.nz;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    int i = c % F.nx + 1, j = (c / F.nx) % F.ny + 1, k = c / (F.nx * F.ny) + 1;
    long q = lidx(F, i, j, k);
    xf[q] = xf[q] + prolong_at(xc, Cc, W, i, j, k);
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void coarse_vcycle_block(const PCArgs<T>& A) {
  extern __shared__ __align__(16) unsigned char pc_smem[];
  T* sm = reinterpret_cast<T*>(pc_smem);
  double* smd = reinterpret_cast<double*>(pc_smem);
  __shared__ double sh[PC_THREADS / 64];
  __shared__ double seq[64];
  const int top = A.nlev - 1;                     // index of the coarsest level
  // b_s from global
  {
    const Lvl<T>& L = A.lv[0];
    int n = L.nx * L.ny * L.nz;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      int i = c % L.nx + 1, j = (c / L.nx) % L.ny + 1, k = c / (L.nx * L.ny) + 1;
      T v = A.gbs[lidx(A.gl, i, j, k)];
      if (top == 0) smd[A.bcoff + c] = (double)v;
      else sm[A.boff[0] + c] = v;
    }
    __syncthreads();
  }
  // down-leg: levels 0 .. top-1 (local numbering)
  for (int q = 0; q < top; q++) {
    const Lvl<T>& L = A.lv[q];
    T* x = sm + A.xoff[q];
    const T* b = sm + A.boff[q];
    pc_smooth(x, b, L, A.omega, A.nu, true);
    if (q + 1 < top) {
      const Lvl<T>& C = A.lv[q + 1];
      pc_restrict<T, T>(sm + A.boff[q + 1], x, b, L, C.nx, C.ny, C.nz, A.rw[q]);
    } else {
      pc_restrict<T, double>(smd + A.bcoff, x, b, L, A.lc.nx, A.lc.ny, A.lc.nz, A.rw[q]);
    }
  }
  // coarsest solve in FP64
  {
    const Lvl<double>& L = A.lc;
    double* x = smd + A.xcoff;
    const double* b = smd + A.bcoff;
    int n = L.nx * L.ny * L.nz;
    pc_zero(x, n);
    if (n == 1) {
      if (threadIdx.x == 0) {
        x[0] = b[0] / diag(L.cwx[1], L.cex[1], L.cwy[1], L.cey[1], L.cwz[1], L.cez[1]);
        if (A.iters_slot) *A.iters_slot = 0;
      }
      __syncthreads();
    } else if (n <= 64) {
      // one wavefront, one cell per lane; LDS traffic of a single wave is
      // ordered, so a wavefront fence replaces the workgroup barriers
      if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const bool own = lane < n;
        const int i = lane % L.nx + 1, j = (lane / L.nx) % L.ny + 1, k = lane / (L.nx * L.ny) + 1;
        const int col = (i + j + k) & 1;     // 1: first sweep (i+j+k odd), 0: second
        auto wsync = []() {
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        };
        auto norm = [&]() -> double {
          double r = own ? resid_at(x, b, L, i, j, k) : 0.0;
          seq[lane] = r * r;
          wsync();
          double s = 0.0;
          if (lane == 0)
            for (int c = 0; c < n; c++) s = s + seq[c];
          s = __shfl(s, 0, 64);
          wsync();
          return s;
        };
        double r0 = norm();
        int it;
        for (it = 0; it < A.maxit; it++) {
          if (own && col == 1) x[lidx(L, i, j, k)] = sor_at(x, b, L, i, j, k, (double)A.omega, x[lidx(L, i, j, k)]);
          wsync();
          if (own && col == 0) x[lidx(L, i, j, k)] = sor_at(x, b, L, i, j, k, (double)A.omega, x[lidx(L, i, j, k)]);
          wsync();
          double rn = norm();
          if (sqrt(rn / r0) <= A.tol) break;
        }
        if (lane == 0 && A.iters_slot) *A.iters_slot = it;
      }
      __syncthreads();
    } else {
      double r0 = coarse_norm(x, b, L, sh, seq);
      int it;
      for (it = 0; it < A.maxit; it++) {
        pc_colour(x, b, L, (double)A.omega, 0, false);
        pc_colour(x, b, L, (double)A.omega, 1, false);
        double rn = coarse_norm(x, b, L, sh, seq);
        if (sqrt(rn / r0) <= A.tol) break;
      }
      if (threadIdx.x == 0 && A.iters_slot) *A.iters_slot = it;
      __syncthreads();
    }
    if (top > 0) {
      T* xt = sm + A.xctoff;
      for (int c = threadIdx.x; c < n; c += blockDim.x) xt[c] = (T)x[c];
      __syncthreads();
    }
  }
  // up-leg
  for (int q = top - 1; q >= 0; q--) {
    const Lvl<T>& L = A.lv[q];
    T* x = sm + A.xoff[q];
    const T* b = sm + A.boff[q];
    if (q + 1 < top) pc_prolong_add(x, sm + A.xoff[q + 1], L, A.lv[q + 1], A.pw[q]);
    else {
      Lvl<T> C = A.lv[top];   // geometry of the coarsest in T, compact
      pc_prolong_add(x, sm + A.xctoff, L, C, A.pw[q]);
    }
    pc_smooth(x, b, L, A.omega, A.nu, false);
  }
  // x_s to global
  {
    const Lvl<T>& L = A.lv[0];
    int n = L.nx * L.ny * L.nz;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      int i = c % L.nx + 1, j = (c / L.nx) % L.ny + 1, k = c / (L.nx * L.ny) + 1;
      A.gxs[lidx(A.gl, i, j, k)] = (top == 0) ? (T)smd[A.xcoff + c] : sm[A.xoff[0] + c];
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(PC_THREADS) k_coarse_vcycle(PCArgs<T> A) {
  coarse_vcycle_block<T>(A);
}

}  // namespace pepsmg
