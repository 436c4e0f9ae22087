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
  auto issue_c = [&](T* dst, int kc) {
    const int kcc = min(max(kc, 1), Cc.nz);
#pragma unroll
    for (int q = 0; q < NCL; q++) {
      int e = min(t + q * NT, CX * CY - 1);
      int cy = e / CX, cx = e - cy * CX;
      int gi = min(max(ic0 - 1 + cx, 1), Cc.nx), gj = min(max(jc0 - 1 + cy, 1), Cc.ny);
      dst[q] = xc[lidx(Cc, gi, gj, kcc)];
    }
  };
  // src: the values issue_c loaded for coarse plane kc
  auto stage_c = [&](const T* src, int kc) {
    const bool vz = kc >= 1 && kc <= Cc.nz;
#pragma unroll
    for (int q = 0; q < NCL; q++) {
      int e = t + q * NT;
      if (e < CX * CY) {
        int cy = e / CX, cx = e - cy * CX;
        int gi = ic0 - 1 + cx, gj = jc0 - 1 + cy;
        bool v = vz && gi >= 1 && gi <= Cc.nx && gj >= 1 && gj <= Cc.ny;
        (&sc[kc & 3][0][0])[e] = v ? src[q] : T(0);
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

  {   // the three staged planes' loads in flight before the first is staged
    T ca[NCL], cb[NCL];
    issue_c(cpre, kc0 - 1);
    issue_c(ca, kc0);
    issue_c(cb, kc0 + 1);
    stage_c(cpre, kc0 - 1);
    stage_c(ca, kc0);
    stage_c(cb, kc0 + 1);
  }
  issue_c(cpre, kc0 + 2);
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
    stage_c(cpre, kc + 2);
    issue_c(cpre, kc + 3);
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
// The 1-D coefficient and transfer-weight arrays of every level are staged
// into LDS at the start, so no phase waits on a global load.
constexpr int PC_MAXLEV = 8;
constexpr int PC_THREADS = 1024;
constexpr int PC_CINV_MAX = 128;   // largest coarsest level solved by its inverse (FP32 inner cycles)

template <typename T>
struct PCArgs {
  int nlev;                       // levels s..L handled: nlev = L - s + 1
  int nu, nu_post;                // smoothing steps on the way down, on the way up
  T omega;
  double tol;
  int maxit;
  Lvl<T> lv[PC_MAXLEV];           // compact LDS geometry (px = nx, pxy = nx*ny)
  Lvl<double> lc;                 // coarsest level, FP64 coefficients
  RestrictW<T> rw[PC_MAXLEV];     // level q -> q+1
  ProlongW<T> pw[PC_MAXLEV];      // level q <- q+1
  int xoff[PC_MAXLEV], boff[PC_MAXLEV];   // LDS offsets (elements of T)
  int coff[PC_MAXLEV];            // staged coefficients and weights of levels 0..nlev-2 (T)
  int roff;                       // residual of the level being restricted (T, level 0's size)
  int xcoff, bcoff, xctoff;       // coarsest: FP64 x, b (in doubles), T copy of x (in T)
  int ccoff;                      // staged coarsest coefficients (doubles)
  Lvl<T> gl;                      // global geometry of level s
  T* gxs;                         // global x_s (output)
  const T* gbs;                   // global b_s (input)
  int* iters_slot;
  // FP32 inner cycles of the mixed-precision solve: the coarsest level's
  // inverse operator, transposed (n x n doubles, entry k n + l = (A^-1)_lk),
  // for an exact coarsest solve in place of the red/black iterations; null
  // otherwise (the FP64 solve keeps the reference's iterations)
  const double* cinv;
};


// Staged arrays of one level in LDS (axis lengths ex = nx+2, ey, ez):
// cwx cex | cwy cey | cwz cez, then for levels with a coarser one the
// restriction weights wx | wy | wz and the prolongation weights
// px nx | py ny | pz nz.  pc_stage_elems gives the length of the block.
__host__ __device__ inline int pc_stage_elems(int nx, int ny, int nz, bool transfer) {
  return (transfer ? 5 : 2) * (nx + ny + nz + 6);
}

// copies level G's arrays (and transfer weights when R, P are given) to s.
// Every load of a thread is issued before its first LDS store (clamped
// indices, no branch next to a load), so staging costs one load latency.
template <typename T>
__device__ void pc_stage(T* s, const Lvl<T>& G, const RestrictW<T>* R, const ProlongW<T>* P) {
  const int ex = G.nx + 2, ey = G.ny + 2, ez = G.nz + 2;
  const int em = max(ex, max(ey, ez));
  for (int c = threadIdx.x; c < em; c += blockDim.x) {
    const int cx = min(c, ex - 1), cy = min(c, ey - 1), cz = min(c, ez - 1);
    T v[15];
    v[0] = G.cwx[cx];
    v[1] = G.cex[cx];
    v[2] = G.cwy[cy];
    v[3] = G.cey[cy];
    v[4] = G.cwz[cz];
    v[5] = G.cez[cz];
    if (R) {
      v[6] = R->wx[cx];
      v[7] = R->wy[cy];
      v[8] = R->wz[cz];
      v[9] = P->px_[cx];
      v[10] = P->nx_[cx];
      v[11] = P->py_[cy];
      v[12] = P->ny_[cy];
      v[13] = P->pz_[cz];
      v[14] = P->nz_[cz];
    }
    if (c < ex) {
      s[c] = v[0];
      s[ex + c] = v[1];
    }
    if (c < ey) {
      s[2 * ex + c] = v[2];
      s[2 * ex + ey + c] = v[3];
    }
    if (c < ez) {
      s[2 * (ex + ey) + c] = v[4];
      s[2 * (ex + ey) + ez + c] = v[5];
    }
    if (R) {
      T* w = s + 2 * (ex + ey + ez);
      T* p = w + (ex + ey + ez);
      if (c < ex) {
        w[c] = v[6];
        p[c] = v[9];
        p[ex + c] = v[10];
      }
      if (c < ey) {
        w[ex + c] = v[7];
        p[2 * ex + c] = v[11];
        p[2 * ex + ey + c] = v[12];
      }
      if (c < ez) {
        w[ex + ey + c] = v[8];
        p[2 * (ex + ey) + c] = v[13];
        p[2 * (ex + ey) + ez + c] = v[14];
      }
    }
  }
}

// level geometry G with its coefficient pointers moved to the staged copy s
template <typename T>
__device__ __forceinline__ Lvl<T> pc_lvl(const Lvl<T>& G, const T* s) {
  const int ex = G.nx + 2, ey = G.ny + 2, ez = G.nz + 2;
  Lvl<T> L = G;
  L.cwx = s;
  L.cex = s + ex;
  L.cwy = s + 2 * ex;
  L.cey = s + 2 * ex + ey;
  L.cwz = s + 2 * (ex + ey);
  L.cez = s + 2 * (ex + ey) + ez;
  return L;
}
template <typename T>
__device__ __forceinline__ RestrictW<T> pc_rw(const RestrictW<T>& W, const Lvl<T>& G, const T* s) {
  const int ex = G.nx + 2, ey = G.ny + 2, ez = G.nz + 2;
  RestrictW<T> R = W;
  const T* w = s + 2 * (ex + ey + ez);
  R.wx = w;
  R.wy = w + ex;
  R.wz = w + ex + ey;
  return R;
}
template <typename T>
__device__ __forceinline__ ProlongW<T> pc_pw(const ProlongW<T>& W, const Lvl<T>& G, const T* s) {
  const int ex = G.nx + 2, ey = G.ny + 2, ez = G.nz + 2;
  ProlongW<T> P = W;
  const T* p = s + 3 * (ex + ey + ez);
  P.px_ = p;
  P.nx_ = p + ex;
  P.py_ = p + 2 * ex;
  P.ny_ = p + 2 * ex + ey;
  P.pz_ = p + 2 * (ex + ey);
  P.nz_ = p + 2 * (ex + ey) + ez;
  return P;
}

// value of lane l of v (l a compile-time lane after unrolling)
__device__ __forceinline__ double pc_readlane(double v, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffu), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Coarsest rbgs_solver on up to 64 cells (src/rbgs_poisson_matrix.f90:47-179)
// by one wavefront, cell c on lane c: coefficients, c0 and b stay in
// registers; x lives in LDS (one wavefront orders its own LDS traffic, so a
// wavefront fence replaces the workgroup barriers).  The norm sums the lanes'
// r^2 in cell order (lanes without a cell add +0.0, which leaves the
// non-negative sum unchanged): the sequential sum of the reference, bit for
// bit.  Returns the iteration count (the reference's loop variable).
static __device__ __forceinline__ int pc_coarsest_wave(double* __restrict__ x, const double* __restrict__ b,
                                             const Lvl<double>& L, double omega, double tol, int maxit) {
  const int n = L.nx * L.ny * L.nz;
  const int lane = threadIdx.x;
  const bool own = lane < n;
  const int c = own ? lane : 0;
  const int i = c % L.nx + 1, j = (c / L.nx) % L.ny + 1, k = c / (L.nx * L.ny) + 1;
  const int col = (i + j + k) & 1;     // 1: first sweep (i+j+k odd), 0: second
  const double c1 = L.cwx[i], c2 = L.cex[i], c3 = L.cwy[j], c4 = L.cey[j], c5 = L.cwz[k], c6 = L.cez[k];
  const double c0 = diag(c1, c2, c3, c4, c5, c6);
  const double bv = b[c];
  const int sy = L.nx, sz = L.nx * L.ny;
  // neighbour cells; -1 outside the level (value 0, the homogeneous ghost)
  const int nw = i > 1 ? c - 1 : -1, ne = i < L.nx ? c + 1 : -1;
  const int ns = j > 1 ? c - sy : -1, nn = j < L.ny ? c + sy : -1;
  const int nb = k > 1 ? c - sz : -1, nt = k < L.nz ? c + sz : -1;
  auto X = [&](int o) -> double { return o >= 0 ? x[o] : 0.0; };
  auto wsync = []() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  auto norm = [&]() -> double {
    double r = 0.0;
    if (own) {
      const double y = c0 * x[c] + c1 * X(nw) + c2 * X(ne) + c3 * X(ns) + c4 * X(nn) + c5 * X(nb) + c6 * X(nt);
      r = bv - y;
    }
    const double r2 = r * r;
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < 64; l++) s = s + pc_readlane(r2, l);
    return s;
  };
  auto sweep = [&](int colour) {
    if (own && col == colour) {
      const double temp = c1 * X(nw) + c2 * X(ne) + c3 * X(ns) + c4 * X(nn) + c5 * X(nb) + c6 * X(nt);
      x[c] = omega * (bv - temp) / c0 + (1.0 - omega) * x[c];
    }
    wsync();
  };
  const double r0 = norm();
  int it;
  for (it = 0; it < maxit; it++) {
    sweep(1);
    sweep(0);
    const double rn = norm();
    if (sqrt(rn / r0) <= tol) break;
  }
  return it;
}

// floor(n / d) for 0 <= n, d < 2^16 (levels held in LDS are far smaller):
// one 64-bit multiply by m = floor((2^32 - 1) / d) + 1 instead of an integer
// division (the error n * (m - 2^32 / d) / 2^32 < n / 2^32 stays below 1/d)
__device__ __forceinline__ unsigned long long pc_magic(int d) { return 0xFFFFFFFFull / (unsigned)d + 1ull; }
__device__ __forceinline__ int pc_div(int n, unsigned long long m) {
  return (int)(((unsigned long long)(unsigned)n * m) >> 32);
}

// Cells of a compact LDS level (px = nx, pxy = nx ny): cell q = (i-1) +
// nx (j-1) + nx ny (k-1); neighbours outside the level read as 0, the
// homogeneous ghost (ldz).  Values and operation order are those of
// sor_at / apply_at.
template <typename T>
struct PCNb {
  T w, e, s, n, b, t;
};
template <typename T>
__device__ __forceinline__ PCNb<T> pc_nb(const T* x, const Lvl<T>& L, int i, int j, int k, int q) {
  const int sy = L.nx, sz = L.nx * L.ny;
  PCNb<T> v;
  v.w = i > 1 ? x[q - 1] : T(0);
  v.e = i < L.nx ? x[q + 1] : T(0);
  v.s = j > 1 ? x[q - sy] : T(0);
  v.n = j < L.ny ? x[q + sy] : T(0);
  v.b = k > 1 ? x[q - sz] : T(0);
  v.t = k < L.nz ? x[q + sz] : T(0);
  return v;
}

template <typename T>
__device__ void pc_colour(T* x, const T* b, const Lvl<T>& L, T omega, int colour, bool zero) {
  const int hx = (L.nx + 1) / 2;
  const int n = hx * L.ny * L.nz;
  const unsigned long long mh = pc_magic(hx), my = pc_magic(L.ny);
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    const int row = pc_div(c, mh);                 // (j-1) + ny (k-1)
    const int tt = c - row * hx, kk = pc_div(row, my);
    const int j = row - kk * L.ny + 1, k = kk + 1;
    const int i = 1 + ((j + k + colour) & 1) + 2 * tt;
    if (i <= L.nx) {
      const int q = (i - 1) + L.nx * row;
      const T c1 = L.cwx[i], c2 = L.cex[i], c3 = L.cwy[j], c4 = L.cey[j], c5 = L.cwz[k], c6 = L.cez[k];
      if (zero) {
        T z = T(0);
        T temp = c1 * z + c2 * z + c3 * z + c4 * z + c5 * z + c6 * z;
        x[q] = omega * (b[q] - temp) / diag(c1, c2, c3, c4, c5, c6) + (T(1) - omega) * z;
      } else {
        const T c0 = diag(c1, c2, c3, c4, c5, c6);
        const PCNb<T> v = pc_nb(x, L, i, j, k, q);
        const T temp = c1 * v.w + c2 * v.e + c3 * v.s + c4 * v.n + c5 * v.b + c6 * v.t;
        x[q] = omega * (b[q] - temp) / c0 + (T(1) - omega) * x[q];
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

// residual of every fine cell into r (compact, one per thread in turn), then
// the restriction from r: the same values and sums as the per-level kernels
template <typename T, typename TC>
__device__ void pc_restrict(TC* bc, T* r, const T* x, const T* b, const Lvl<T>& F, int cnx, int cny, int cnz,
                            const RestrictW<T>& W) {
  const int nf = F.nx * F.ny * F.nz;
  const unsigned long long mx = pc_magic(F.nx), my = pc_magic(F.ny);
  for (int c = threadIdx.x; c < nf; c += blockDim.x) {
    const int row = pc_div(c, mx), kk = pc_div(row, my);
    const int i = c - row * F.nx + 1, j = row - kk * F.ny + 1, k = kk + 1;
    const T c1 = F.cwx[i], c2 = F.cex[i], c3 = F.cwy[j], c4 = F.cey[j], c5 = F.cwz[k], c6 = F.cez[k];
    const T c0 = diag(c1, c2, c3, c4, c5, c6);
    const PCNb<T> v = pc_nb(x, F, i, j, k, c);
    const T y = c0 * x[c] + c1 * v.w + c2 * v.e + c3 * v.s + c4 * v.n + c5 * v.b + c6 * v.t;
    r[c] = b[c] - y;
  }
  __syncthreads();
  const int n = cnx * cny * cnz;
  const unsigned long long mcx = pc_magic(cnx), mcy = pc_magic(cny);
  const int sy = F.nx, sz = F.nx * F.ny;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    const int row = pc_div(c, mcx), kk = pc_div(row, mcy);
    const int ic = c - row * cnx + 1, jc = row - kk * cny + 1, kc = kk + 1;
    int iz, ip, jz, jp, kz, kp;
    T ax1, ax2, ay1, ay2, az1, az2;
    restrict_axis(W.wx, W.sx, ic, iz, ip, ax1, ax2);
    restrict_axis(W.wy, W.sy, jc, jz, jp, ay1, ay2);
    restrict_axis(W.wz, W.sz, kc, kz, kp, az1, az2);
    T v111 = ax1 * ay1 * az1, v211 = ax2 * ay1 * az1, v121 = ax1 * ay2 * az1, v221 = ax2 * ay2 * az1;
    T v112 = ax1 * ay1 * az2, v212 = ax2 * ay1 * az2, v122 = ax1 * ay2 * az2, v222 = ax2 * ay2 * az2;
    T vol_c = v111 + v211 + v121 + v221 + v112 + v212 + v122 + v222;
    auto R = [&](int a, int bb, int cc) { return r[(a - 1) + sy * (bb - 1) + sz * (cc - 1)]; };
    T s = v111 * R(ip, jp, kp) + v211 * R(iz, jp, kp) + v121 * R(ip, jz, kp) + v221 * R(iz, jz, kp) +
          v112 * R(ip, jp, kz) + v212 * R(iz, jp, kz) + v122 * R(ip, jz, kz) + v222 * R(iz, jz, kz);
    bc[c] = (TC)(s / vol_c);
  }
  __syncthreads();
}

// x_f += P x_c (prolong_at's weights, terms and order) on a compact level
template <typename T>
__device__ void pc_prolong_add(T* xf, const T* xc, const Lvl<T>& F, const Lvl<T>& Cc, const ProlongW<T>& W) {
  const int n = F.nx * F.ny * F.nz;
  const unsigned long long mx = pc_magic(F.nx), my = pc_magic(F.ny);
  auto V = [&](int a, int bb, int cc) -> T {
    return (a >= 1 && a <= Cc.nx && bb >= 1 && bb <= Cc.ny && cc >= 1 && cc <= Cc.nz)
               ? xc[(a - 1) + Cc.nx * ((bb - 1) + Cc.ny * (cc - 1))]
               : T(0);
  };
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    const int row = pc_div(c, mx), kk = pc_div(row, my);
    const int i = c - row * F.nx + 1, j = row - kk * F.ny + 1, k = kk + 1;
    int ci, ni, cj, nj, ck, nk;
    T wsx, wnx, wsy, wny, wsz, wnz;
    prolong_axis(W.px_, W.nx_, W.sx, i, ci, ni, wsx, wnx);
    prolong_axis(W.py_, W.ny_, W.sy, j, cj, nj, wsy, wny);
    prolong_axis(W.pz_, W.nz_, W.sz, k, ck, nk, wsz, wnz);
    T num = wsx * wsy * wsz * V(ci, cj, ck) + wnx * wsy * wsz * V(ni, cj, ck) + wsx * wny * wsz * V(ci, nj, ck) +
            wsx * wsy * wnz * V(ci, cj, nk) + wnx * wny * wsz * V(ni, nj, ck) + wnx * wsy * wnz * V(ni, cj, nk) +
            wsx * wny * wnz * V(ci, nj, nk) + wnx * wny * wnz * V(ni, nj, nk);
    xf[c] = xf[c] + num / ((wsx + wnx) * (wsy + wny) * (wsz + wnz));
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
  // coefficients and transfer weights into LDS, b_s from global
  for (int q = 0; q < top; q++) pc_stage(sm + A.coff[q], A.lv[q], &A.rw[q], &A.pw[q]);
  pc_stage(smd + A.ccoff, A.lc, (const RestrictW<double>*)nullptr, (const ProlongW<double>*)nullptr);
  {
    // four loads in flight per thread (clamped cells) before the LDS stores
    const Lvl<T>& L = A.lv[0];
    const int n = L.nx * L.ny * L.nz;
    const int bd = blockDim.x;
    for (int c0 = threadIdx.x; c0 < n; c0 += 4 * bd) {
      T v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int c = min(c0 + u * bd, n - 1);
        const int i = c % L.nx + 1, j = (c / L.nx) % L.ny + 1, k = c / (L.nx * L.ny) + 1;
        v[u] = A.gbs[lidx(A.gl, i, j, k)];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int c = c0 + u * bd;
        if (c < n) {
          if (top == 0) smd[A.bcoff + c] = (double)v[u];
          else sm[A.boff[0] + c] = v[u];
        }
      }
    }
    __syncthreads();
  }
  // down-leg: levels 0 .. top-1 (local numbering)
  for (int q = 0; q < top; q++) {
    const T* st = sm + A.coff[q];
    const Lvl<T> L = pc_lvl(A.lv[q], st);
    const RestrictW<T> W = pc_rw(A.rw[q], A.lv[q], st);
    T* x = sm + A.xoff[q];
    const T* b = sm + A.boff[q];
    pc_smooth(x, b, L, A.omega, A.nu, true);
    if (q + 1 < top) {
      const Lvl<T>& C = A.lv[q + 1];
      pc_restrict<T, T>(sm + A.boff[q + 1], sm + A.roff, x, b, L, C.nx, C.ny, C.nz, W);
    } else {
      pc_restrict<T, double>(smd + A.bcoff, sm + A.roff, x, b, L, A.lc.nx, A.lc.ny, A.lc.nz, W);
    }
  }
  // coarsest solve in FP64
  {
    const Lvl<double> L = pc_lvl(A.lc, (const double*)(smd + A.ccoff));
    double* x = smd + A.xcoff;
    const double* b = smd + A.bcoff;
    int n = L.nx * L.ny * L.nz;
    pc_zero(x, n);
    if (A.cinv && n <= PC_CINV_MAX) {   // x = A^-1 b, one lane per cell
      if ((int)threadIdx.x < n) {
        double v = 0.0;
#pragma unroll 8
        for (int k = 0; k < n; k++) v += A.cinv[(long)k * n + threadIdx.x] * b[k];
        x[threadIdx.x] = v;
      }
      if (threadIdx.x == 0 && A.iters_slot) *A.iters_slot = 0;
      __syncthreads();
    } else if (n == 1) {
      if (threadIdx.x == 0) {
        x[0] = b[0] / diag(L.cwx[1], L.cex[1], L.cwy[1], L.cey[1], L.cwz[1], L.cez[1]);
        if (A.iters_slot) *A.iters_slot = 0;
      }
      __syncthreads();
    } else if (n <= 64) {
      if (threadIdx.x < 64) {
        const int it = pc_coarsest_wave(x, b, L, (double)A.omega, A.tol, A.maxit);
        if (threadIdx.x == 0 && A.iters_slot) *A.iters_slot = it;
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
    const T* st = sm + A.coff[q];
    const Lvl<T> L = pc_lvl(A.lv[q], st);
    const ProlongW<T> P = pc_pw(A.pw[q], A.lv[q], st);
    T* x = sm + A.xoff[q];
    const T* b = sm + A.boff[q];
    if (q + 1 < top) pc_prolong_add(x, sm + A.xoff[q + 1], L, A.lv[q + 1], P);
    else {
      Lvl<T> C = A.lv[top];   // geometry of the coarsest in T, compact
      pc_prolong_add(x, sm + A.xctoff, L, C, P);
    }
    pc_smooth(x, b, L, A.omega, A.nu_post, false);
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
