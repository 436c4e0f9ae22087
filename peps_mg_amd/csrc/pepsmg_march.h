// pepsmg_march.h -- residual kernels that march through z planes (gfx950).
//
// All of them evaluate r = b - A x exactly as multigrid_residual /
// mv_mul_poisson_matrix do (src/multigrid.f90:1795-1805,
// src/poisson_matrix_operator.f90:150-156), with x staged plane by plane in
// LDS (a 64 x TY tile with a ring of 2 cells in x, so rows load as aligned
// 16-byte pairs, and 1 cell in y), and the next plane's loads in flight in
// registers while the current plane is computed.  Each element of x and b
// is read from HBM once; ring cells are re-read from L2.  Blocks are
// numbered with the z-chunk fastest (one z-slab per XCD when there are 8 or
// more chunks).
//
//   k_resid_march<T,TY,0>: partial sums of r^2 (convergence check,
//                          src/multigrid.f90:1970-1971)
//   k_resid_march<T,TY,1>: same, and stores r
//   k_resid_march<double,TY,2>: iterative-refinement step: x_in = x + e
//       (e the FP32 correction), stores x_in, r as float, partial r^2; x
//       and b in layouts of their own (the solver's, or the caller's padded
//       arrays, whose pairs are only 8-byte aligned)
//   k_cg_ap_fold<T,TY> (KIND 7): KIND 3 on pnew = r + beta p, formed while
//                          staging; pnew stored too (the direction update
//                          of the previous CG iteration, folded in)
//   k_resid_march<T,TY,3>: CG step Ap = A p (stored in rout) and partial sums
//                          of (Ap, p) (src/cg_poisson_matrix.f90:206-207)
//   k_resid_march<T,TY,4>: CG start r = b - A x stored in rout and in xnew (p),
//                          partial sums of (r, r) (:196-204)
//   k_resid_march<double,TY,6>: the refinement step expected to be the last:
//       x_in = x + e stored into the caller's padded array only (no copy-out
//       pass, no solver-layout x, no r32) and partial r^2
//   k_refine_start:        first iterative-refinement step read straight
//       from the caller's padded (0:n+1)^3 arrays (unaligned rows): stores
//       only r as float and partial r^2 (x stays in the caller's array until
//       the first update, b for the whole solve)
//   k_restrict_march<T>:   residual + restriction into the next level's b
//       (src/multigrid.f90:1932-1933, :1939-1940 with :1460-1494), for
//       levels that halve every axis
#pragma once
#include "pepsmg_fused.h"

namespace pepsmg {

constexpr int MARCH_THREADS = 256;

struct MarchGrid {
  int gx, gy, nch, kchunk;
};

// Stages x planes (4 slots) of a TX x TY tile: x = i0-2 .. i0+TX+1, y = j0-1 .. j0+TY.
// issue() loads one plane into registers from clamped in-bounds addresses
// with no branch around the loads; stage() zeroes what lies outside the
// domain, adds the FP32 correction (MIXED) or forms the CG direction
// x + beta * cgp (CGP: x is r, cgp the previous direction) and writes LDS.
// Keeping every select away from the loads lets the compiler wait only for
// the loads a plane needs instead of draining the ones in flight.
template <typename T, int TY, bool MIXED, bool UNAL = false, int TX = 64, int NT = MARCH_THREADS, bool CGP = false>
struct XStager {
  static constexpr int RX = TX + 4, RY = TY + 2, XP = RX / 2;
  static constexpr int NXL = (RY * XP + NT - 1) / NT;
  using V2 = typename Vec2<T>::type;
  V2 pre[NXL];
  float2 epre[MIXED ? NXL : 1];
  V2 ppre[CGP ? NXL : 1];
  const T* cgp = nullptr;
  T beta = T(0);
  int mq;

  // LX: layout of x; LE: layout of e and cgp
  __device__ __forceinline__ void issue(const T* __restrict__ x, const Lvl<T>& LX, const float* __restrict__ e,
                                        const Lvl<T>& LE, int i0, int j0, int m) {
    mq = m;
    const int mc = min(max(m, 1), LE.nz);
#pragma unroll
    for (int q = 0; q < NXL; q++) {
      int idx = min((int)threadIdx.x + q * NT, RY * XP - 1);
      int ry = idx / XP, px = idx - ry * XP;
      int gi = i0 - 2 + 2 * px, gj = j0 - 1 + ry;      // gi is odd
      // an odd gi <= nx keeps the 2-element load inside the padded row
      int gic = (gi >= 1 && gi <= LE.nx) ? gi : 1;
      int gjc = min(max(gj, 1), LE.ny);
      if (UNAL) {   // x in a layout of its own, pairs possibly only 8-byte aligned
        pre[q] = ld_pair_unal(x + lidx(LX, gic, gjc, mc));
      } else {
        pre[q] = *reinterpret_cast<const V2*>(x + lidx(LX, gic, gjc, mc));
      }
      const long qe = lidx(LE, gic, gjc, mc);
      if (MIXED) epre[q] = *reinterpret_cast<const float2*>(e + qe);
      if (CGP) ppre[q] = *reinterpret_cast<const V2*>(cgp + qe);
    }
  }
  __device__ __forceinline__ void stage(T (*s)[RX], const Lvl<T>& L, int i0, int j0) {
    const bool vz = mq >= 1 && mq <= L.nz;
#pragma unroll
    for (int q = 0; q < NXL; q++) {
      int idx = threadIdx.x + q * NT;
      if (idx < RY * XP) {
        int ry = idx / XP, px = idx - ry * XP;
        int gi = i0 - 2 + 2 * px, gj = j0 - 1 + ry;
        bool vrow = vz && gj >= 1 && gj <= L.ny;
        bool va = vrow && gi >= 1 && gi <= L.nx;
        bool vb = vrow && gi + 1 >= 1 && gi + 1 <= L.nx;
        T a = pre[q].x, b = pre[q].y;
        if (MIXED) {
          a = a + (T)epre[q].x;
          b = b + (T)epre[q].y;
        }
        if (CGP) {   // p = r + beta p (src/cg_poisson_matrix.f90:229)
          a = a + beta * ppre[q].x;
          b = b + beta * ppre[q].y;
        }
        s[ry][2 * px] = va ? a : T(0);
        s[ry][2 * px + 1] = vb ? b : T(0);
      }
    }
  }
};

// r at tile cell (lx, ly) of plane slots (s0 = m, sm = m-1, sp = m+1);
// LDS column of cell lx is lx+2, row ly+1.
// (A x) at tile cell (lx, ly), in mv_mul's order (c0 term first)
template <typename T, int RX>
__device__ __forceinline__ T march_apply(T (*s0)[RX], T (*sm)[RX], T (*sp)[RX], int lx, int ly, T c1, T c2, T c3,
                                         T c4, T c5, T c6) {
  T c0 = diag(c1, c2, c3, c4, c5, c6);
  int cx = lx + 2, cy = ly + 1;
  return c0 * s0[cy][cx] + c1 * s0[cy][cx - 1] + c2 * s0[cy][cx + 1] + c3 * s0[cy - 1][cx] + c4 * s0[cy + 1][cx] +
         c5 * sm[cy][cx] + c6 * sp[cy][cx];
}

template <typename T, int RX>
__device__ __forceinline__ T march_resid(T (*s0)[RX], T (*sm)[RX], T (*sp)[RX], int lx, int ly, T c1, T c2, T c3,
                                         T c4, T c5, T c6, T bv) {
  return bv - march_apply<T, RX>(s0, sm, sp, lx, ly, c1, c2, c3, c4, c5, c6);
}

// L: the solver's layout (tile geometry; e, rout, r32, xnew).  LX, LB:
// layouts of the inputs x and b for the refinement kinds 2, 5 and 6 (the
// solver's, or the caller's padded arrays, read as 8-byte-aligned pairs);
// the other kinds read both in L.  uout, LU: KIND 6's output, the caller's
// padded array.
template <typename T, int TY, int KIND>
__device__ __forceinline__ void resid_march_block(int bid, const T* __restrict__ x, const float* __restrict__ e,
                                                  const T* __restrict__ b, T* __restrict__ rout,
                                                  float* __restrict__ r32, T* __restrict__ xnew, const Lvl<T>& L,
                                                  const MarchGrid& G, double* __restrict__ partial,
                                                  const Lvl<T>& LX, const Lvl<T>& LB, T* __restrict__ uout,
                                                  const Lvl<T>& LU, T beta = T(0)) {
  constexpr bool MIXED = (KIND == 2 || KIND == 6);
  constexpr bool UNAL = (KIND == 2 || KIND == 5);            // x in layout LX
  constexpr bool BU = (KIND == 2 || KIND == 5 || KIND == 6);  // b in layout LB
  constexpr bool CGP = (KIND == 7);   // x is r, b the previous direction p
  constexpr bool AP = (KIND == 3 || KIND == 7);
  using St = XStager<T, TY, MIXED, UNAL, 64, MARCH_THREADS, CGP>;
  constexpr int RX = St::RX, RY = St::RY;
  constexpr int NP = (TY * 32 + MARCH_THREADS - 1) / MARCH_THREADS;   // cell pairs per thread
  using V2 = typename Vec2<T>::type;
  __shared__ T sx[4][RY][RX];
  __shared__ double sh[MARCH_THREADS / 64];

  const int bz = bid % G.nch;
  const int rxy = bid / G.nch;
  const int i0 = 1 + (rxy % G.gx) * 64;
  const int j0 = 1 + (rxy / G.gx) * TY;
  const int k0 = 1 + bz * G.kchunk;
  const int k1 = min(k0 + G.kchunk, L.nz + 1);

  // per-thread pairs: (ly, lp) -> cells i0+2lp, i0+2lp+1 of row j0+ly
  T c1a[NP], c2a[NP], c1b[NP], c2b[NP], c3[NP], c4[NP];
  int ly[NP], lp[NP];
  bool act[NP];
#pragma unroll
  for (int q = 0; q < NP; q++) {
    int idx = threadIdx.x + q * MARCH_THREADS;
    ly[q] = idx / 32;
    lp[q] = idx - ly[q] * 32;
    int gi = i0 + 2 * lp[q], gj = j0 + ly[q];
    act[q] = idx < TY * 32 && gj <= L.ny && gi <= L.nx;
    c1a[q] = act[q] ? L.cwx[gi] : T(0);
    c2a[q] = act[q] ? L.cex[gi] : T(0);
    c1b[q] = (act[q] && gi + 1 <= L.nx) ? L.cwx[gi + 1] : T(0);
    c2b[q] = (act[q] && gi + 1 <= L.nx) ? L.cex[gi + 1] : T(0);
    c3[q] = act[q] ? L.cwy[gj] : T(0);
    c4[q] = act[q] ? L.cey[gj] : T(0);
  }

  St st;
  V2 bpre[NP];
  // b of this thread's pairs, from clamped addresses (values of cells
  // outside the domain are never used)
  auto issue_b = [&](int m) {
    if (AP) return;                          // A p needs no right-hand side
    const int mc = min(max(m, 1), L.nz);
#pragma unroll
    for (int q = 0; q < NP; q++) {
      int gi = act[q] ? i0 + 2 * lp[q] : 1, gj = act[q] ? j0 + ly[q] : 1;
      if (BU) {
        bpre[q] = ld_pair_unal(b + lidx(LB, gi, gj, mc));
      } else {
        bpre[q] = ld_stream_v2(b + lidx(L, gi, gj, mc));
      }
    }
  };

  const Lvl<T>& LXI = UNAL ? LX : L;   // layout of the x input
  if (CGP) {
    st.cgp = b;
    st.beta = beta;
  }
  {   // both staged planes' loads in flight before the first is staged
    St s2;
    if (CGP) {
      s2.cgp = b;
      s2.beta = beta;
    }
    st.issue(x, LXI, e, L, i0, j0, k0 - 1);
    s2.issue(x, LXI, e, L, i0, j0, k0);
    st.stage(sx[(k0 - 1) & 3], L, i0, j0);
    s2.stage(sx[k0 & 3], L, i0, j0);
  }
  st.issue(x, LXI, e, L, i0, j0, k0 + 1);
  issue_b(k0);
  double acc = 0.0;
  // plane step; the loop below runs it twice per iteration so that the
  // loads in flight across the back edge sit in the same registers as the
  // prologue's and the compiler's wait counts stay exact
  auto body = [&](int m) {
    st.stage(sx[(m + 1) & 3], L, i0, j0);
    lds_barrier();
    T(*s0)[RX] = sx[m & 3];
    T(*sm)[RX] = sx[(m - 1) & 3];
    T(*sp)[RX] = sx[(m + 1) & 3];
    T c5 = ldc(L.cwz, m), c6 = ldc(L.cez, m);
    V2 bv[NP];
#pragma unroll
    for (int q = 0; q < NP; q++) bv[q] = bpre[q];
    st.issue(x, LXI, e, L, i0, j0, m + 2);
    issue_b(m + 1);
#pragma unroll
    for (int q = 0; q < NP; q++) {
      if (!act[q]) continue;
      int gi = i0 + 2 * lp[q], gj = j0 + ly[q];
      bool two = gi + 1 <= L.nx;
      T ra, rb;
      if (AP) {
        ra = march_apply<T, RX>(s0, sm, sp, 2 * lp[q], ly[q], c1a[q], c2a[q], c3[q], c4[q], c5, c6);
        rb = two ? march_apply<T, RX>(s0, sm, sp, 2 * lp[q] + 1, ly[q], c1b[q], c2b[q], c3[q], c4[q], c5, c6) : T(0);
        acc += (double)ra * (double)s0[ly[q] + 1][2 * lp[q] + 2];
        if (two) acc += (double)rb * (double)s0[ly[q] + 1][2 * lp[q] + 3];
      } else {
        ra = march_resid<T, RX>(s0, sm, sp, 2 * lp[q], ly[q], c1a[q], c2a[q], c3[q], c4[q], c5, c6, bv[q].x);
        rb = two ? march_resid<T, RX>(s0, sm, sp, 2 * lp[q] + 1, ly[q], c1b[q], c2b[q], c3[q], c4[q], c5, c6, bv[q].y)
                 : T(0);
        acc += (double)ra * (double)ra;
        acc += (double)rb * (double)rb;
      }
      long q0 = lidx(L, gi, gj, m);
      if (KIND == 1 || AP || KIND == 4) {
        if (two) st_stream_pair(rout + q0, ra, rb);
        else rout[q0] = ra;
      }
      if (KIND == 7) {   // the new direction p of this tile's cells
        if (two) st_stream_pair(xnew + q0, s0[ly[q] + 1][2 * lp[q] + 2], s0[ly[q] + 1][2 * lp[q] + 3]);
        else xnew[q0] = s0[ly[q] + 1][2 * lp[q] + 2];
      }
      if (KIND == 4) {
        if (two) st_stream_pair(xnew + q0, ra, rb);
        else xnew[q0] = ra;
      }
      if (KIND == 5) {   // the residual only: x and b stay in the caller's arrays
        if (two) st_stream_pair(r32 + q0, (float)ra, (float)rb);
        else r32[q0] = (float)ra;
      }
      if (KIND == 6) {   // x_in into the caller's padded array (layout LU)
        const long u = lidx(LU, gi, gj, m);
        if (two) st_pair_unal(uout + u, s0[ly[q] + 1][2 * lp[q] + 2], s0[ly[q] + 1][2 * lp[q] + 3]);
        else uout[u] = s0[ly[q] + 1][2 * lp[q] + 2];
      }
      if (KIND == 2) {
        if (two) {
          V2 w;
          w.x = s0[ly[q] + 1][2 * lp[q] + 2];
          w.y = s0[ly[q] + 1][2 * lp[q] + 3];
          st_stream_pair(xnew + q0, w.x, w.y);
          float2 f;
          f.x = (float)ra;
          f.y = (float)rb;
          st_stream_pair(r32 + q0, f.x, f.y);
        } else {
          xnew[q0] = s0[ly[q] + 1][2 * lp[q] + 2];
          r32[q0] = (float)ra;
        }
      }
    }
  };
  int m = k0;
  for (; m + 1 < k1; m += 2) {
    body(m);
    body(m + 1);
  }
  if (m < k1) body(m);
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[bid] = s;
}

template <typename T, int TY, int KIND>
__global__ void __launch_bounds__(MARCH_THREADS)
k_resid_march(const T* __restrict__ x, const float* __restrict__ e, const T* __restrict__ b, T* __restrict__ rout,
              float* __restrict__ r32, T* __restrict__ xnew, Lvl<T> L, MarchGrid G, double* __restrict__ partial) {
  resid_march_block<T, TY, KIND>(blockIdx.x, x, e, b, rout, r32, xnew, L, G, partial, L, L, (T*)nullptr, L);
}

// Refinement step (KIND 2): x + e into xnew (layout L) and r32, partial sums
// of r^2; x in layout LX and b in layout LB (the solver's, or the caller's
// padded arrays: pitches nx+2 and (nx+2)(ny+2), base at cell (1,1,1)).
template <int TY>
__global__ void __launch_bounds__(MARCH_THREADS)
k_refine_step(const double* __restrict__ x, const float* __restrict__ e, const double* __restrict__ b,
              float* __restrict__ r32, double* __restrict__ xnew, Lvl<double> L, Lvl<double> LX, Lvl<double> LB,
              MarchGrid G, double* __restrict__ partial) {
  resid_march_block<double, TY, 2>(blockIdx.x, x, e, b, (double*)nullptr, r32, xnew, L, G, partial, LX, LB,
                                   (double*)nullptr, L);
}

// CG step with the direction update folded in (KIND 7): pnew = r + beta p
// (beta = sc[new_slot] / sc[old_slot], src/cg_poisson_matrix.f90:223-229)
// formed while staging, stored for the tile cells, Ap = A pnew stored in
// ap, partial sums of (Ap, pnew) (:206-207).  p and pnew are different
// arrays (neighbouring tiles still read p).
template <typename T, int TY>
__global__ void __launch_bounds__(MARCH_THREADS)
k_cg_ap_fold(const T* __restrict__ r, const T* __restrict__ p, T* __restrict__ pnew, T* __restrict__ ap, Lvl<T> L,
             MarchGrid G, double* __restrict__ partial, const double* __restrict__ sc, int new_slot, int old_slot) {
  const T beta = (T)(sc[new_slot] / sc[old_slot]);
  resid_march_block<T, TY, 7>(blockIdx.x, r, (const float*)nullptr, p, ap, (float*)nullptr, pnew, L, G, partial, L, L,
                              (T*)nullptr, L, beta);
}

// Refinement step expected to be the last (KIND 6): x_in = x + e goes only
// into the caller's padded array usr (layout U: pitches nx+2 and
// (nx+2)(ny+2), base at cell (1,1,1)), with the partial sums of r^2; x and e
// in the solver's layout L, b in layout LB.  No r32, no solver-layout x.
template <int TY>
__global__ void __launch_bounds__(MARCH_THREADS)
k_refine_last(const double* __restrict__ x, const float* __restrict__ e, const double* __restrict__ b, Lvl<double> L,
              Lvl<double> LB, Lvl<double> U, MarchGrid G, double* __restrict__ partial, double* __restrict__ usr) {
  resid_march_block<double, TY, 6>(blockIdx.x, x, e, b, (double*)nullptr, (float*)nullptr, (double*)nullptr, L, G,
                                   partial, L, LB, usr, U);
}

// First refinement step from the caller's arrays (layout U of x and b):
// r32 = b - A x and the partial sums of r^2, nothing else.
template <int TY>
__global__ void __launch_bounds__(MARCH_THREADS)
k_refine_start(const double* __restrict__ x, const double* __restrict__ b, float* __restrict__ r32, Lvl<double> L,
               Lvl<double> U, MarchGrid G, double* __restrict__ partial) {
  resid_march_block<double, TY, 5>(blockIdx.x, x, (const float*)nullptr, b, (double*)nullptr, r32, (double*)nullptr, L,
                                   G, partial, U, U, (double*)nullptr, U);
}

// Residual + restriction for a level that halves every axis.  Fine tile
// TX x 16 (TX = 64, or 128 for FP32: the same bytes per row), coarse tile
// TX/2 x 8, one coarse cell per thread (4 TX threads); chunks hold an even
// number of fine planes.  Each thread evaluates the residuals of its coarse
// cell's own children: the fine cell pair under it on the two fine rows of
// each of the two fine planes.  The four residuals of the first (odd) plane
// wait in registers for the second, so the eight-term sum needs no LDS
// round trip and no second barrier.
template <int TX>
constexpr int march_threads() { return 4 * TX; }

template <typename T, int TX>
__device__ __forceinline__ void restrict_march_block(int bid, T* __restrict__ bc, const T* __restrict__ x,
                                                     const T* __restrict__ b, const Lvl<T>& L, const Lvl<T>& Cc,
                                                     const RestrictW<T>& W, const MarchGrid& G) {
  constexpr int TY = 16, TP = TX / 2, NT = march_threads<TX>();
  using St = XStager<T, TY, false, false, TX, NT>;
  constexpr int RX = St::RX;
  using V2 = typename Vec2<T>::type;
  __shared__ T sx[4][St::RY][RX];

  const int bz = bid % G.nch;
  const int rxy = bid / G.nch;
  const int i0 = 1 + (rxy % G.gx) * TX;
  const int j0 = 1 + (rxy / G.gx) * TY;
  const int k0 = 1 + bz * G.kchunk;                 // odd
  const int k1 = min(k0 + G.kchunk, L.nz + 1);

  // coarse cell of this thread: (ic, jc) = (ic0 + t%TP, jc0 + t/TP); its
  // children are fine cells iz = 2ic-1, ip = 2ic (tile columns 2 lic and
  // 2 lic + 1) on rows jz = 2jc-1, jp = 2jc (tile rows 2 ljc and 2 ljc + 1)
  const int t = threadIdx.x;
  const int lic = t % TP, ljc = t / TP;
  const int ic = (i0 + 1) / 2 + lic, jc = (j0 + 1) / 2 + ljc;
  const bool cact = ic <= Cc.nx && jc <= Cc.ny;
  const int iz = 2 * ic - 1, ip = 2 * ic, jz = 2 * jc - 1, jp = 2 * jc;
  const int lxz = 2 * lic, lyz = 2 * ljc;
  T c1a = T(0), c2a = T(0), c1b = T(0), c2b = T(0), c3z = T(0), c4z = T(0), c3p = T(0), c4p = T(0);
  T ax1 = T(0), ax2 = T(0), ay1 = T(0), ay2 = T(0);
  if (cact) {   // a halving level: every child of an active coarse cell is in the domain
    c1a = L.cwx[iz]; c2a = L.cex[iz]; c1b = L.cwx[ip]; c2b = L.cex[ip];
    c3z = L.cwy[jz]; c4z = L.cey[jz]; c3p = L.cwy[jp]; c4p = L.cey[jp];
    ax1 = W.wx[iz]; ax2 = W.wx[ip]; ay1 = W.wy[jz]; ay2 = W.wy[jp];
  }

  St st;
  V2 bpre[2];   // b of the children pairs on rows jz and jp
  // from clamped addresses (values of cells outside the domain are never used)
  auto issue_b = [&](int m) {
    const int mc = min(max(m, 1), L.nz);
    const int gi = cact ? iz : 1, g0 = cact ? jz : 1, g1 = cact ? jp : 1;
    bpre[0] = ld_stream_v2(b + lidx(L, gi, g0, mc));
    bpre[1] = ld_stream_v2(b + lidx(L, gi, g1, mc));
  };

  {   // both staged planes' loads in flight before the first is staged
    St s2;
    st.issue(x, L, nullptr, L, i0, j0, k0 - 1);
    s2.issue(x, L, nullptr, L, i0, j0, k0);
    st.stage(sx[(k0 - 1) & 3], L, i0, j0);
    s2.stage(sx[k0 & 3], L, i0, j0);
  }
  st.issue(x, L, nullptr, L, i0, j0, k0 + 1);
  issue_b(k0);
  // residuals of the odd plane kz = 2kc-1: (x child, y child) = (z,z), (p,z), (z,p), (p,p)
  T rzz = T(0), rpz = T(0), rzp = T(0), rpp = T(0);
  // plane step, run twice per iteration (chunks hold an even number of
  // planes, starting on an odd one) so the compiler's wait counts across
  // the back edge stay exact and each step knows its plane's parity
  auto step = [&](int m, auto odd_t) {
    constexpr bool ODD = decltype(odd_t)::value;
    st.stage(sx[(m + 1) & 3], L, i0, j0);
    lds_barrier();
    T(*s0)[RX] = sx[m & 3];
    T(*sm)[RX] = sx[(m - 1) & 3];
    T(*sp)[RX] = sx[(m + 1) & 3];
    T c5 = ldc(L.cwz, m), c6 = ldc(L.cez, m);
    V2 bz0 = bpre[0], bz1 = bpre[1];
    st.issue(x, L, nullptr, L, i0, j0, m + 2);
    issue_b(m + 1);
    T qzz = T(0), qpz = T(0), qzp = T(0), qpp = T(0);
    if (cact) {
      qzz = march_resid<T, RX>(s0, sm, sp, lxz, lyz, c1a, c2a, c3z, c4z, c5, c6, bz0.x);
      qpz = march_resid<T, RX>(s0, sm, sp, lxz + 1, lyz, c1b, c2b, c3z, c4z, c5, c6, bz0.y);
      qzp = march_resid<T, RX>(s0, sm, sp, lxz, lyz + 1, c1a, c2a, c3p, c4p, c5, c6, bz1.x);
      qpp = march_resid<T, RX>(s0, sm, sp, lxz + 1, lyz + 1, c1b, c2b, c3p, c4p, c5, c6, bz1.y);
    }
    if (ODD) {
      rzz = qzz; rpz = qpz; rzp = qzp; rpp = qpp;
    } else if (cact) {
      // (src/multigrid.f90:1460-1494) the even plane kp = 2kc's four terms
      // first, then the odd plane's, each in the order (p,p), (z,p), (p,z), (z,z)
      const int kc = m / 2;
      T az1 = ldc(W.wz, m - 1), az2 = ldc(W.wz, m);
      T v111 = ax1 * ay1 * az1, v211 = ax2 * ay1 * az1, v121 = ax1 * ay2 * az1, v221 = ax2 * ay2 * az1;
      T v112 = ax1 * ay1 * az2, v212 = ax2 * ay1 * az2, v122 = ax1 * ay2 * az2, v222 = ax2 * ay2 * az2;
      T vol_c = v111 + v211 + v121 + v221 + v112 + v212 + v122 + v222;
      T s = v111 * qpp + v211 * qzp + v121 * qpz + v221 * qzz + v112 * rpp + v212 * rzp + v122 * rpz + v222 * rzz;
      __builtin_nontemporal_store(s / vol_c, bc + lidx(Cc, ic, jc, kc));
    }
  };
  for (int m = k0; m < k1; m += 2) {
    step(m, std::true_type{});
    step(m + 1, std::false_type{});
  }
}

template <typename T, int TX>
__global__ void __launch_bounds__(march_threads<TX>())
k_restrict_march(T* __restrict__ bc, const T* __restrict__ x, const T* __restrict__ b, Lvl<T> L, Lvl<T> Cc,
                 RestrictW<T> W, MarchGrid G) {
  restrict_march_block<T, TX>(blockIdx.x, bc, x, b, L, Cc, W, G);
}

}  // namespace pepsmg
