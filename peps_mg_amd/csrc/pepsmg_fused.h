// pepsmg_fused.h -- single-pass red/black iteration (gfx950).
//
// One launch performs one full red+black iteration of
// rbgs_iterator_poisson_matrix (src/rbgs_poisson_matrix.f90:239-271),
// out of place: x_in is read once, b once, x_out written once (3 n elements,
// against 6 n for two in-place colour sweeps).  Red cells only read black
// neighbours of x_in; black cells read the new red values, which each block
// computes itself for its tile plus a one-cell ring, so blocks never read
// each other's output and need no inter-block ordering.  Every value is the
// one the two in-place sweeps produce (same expression and operand order).
//
// Blocking: a block owns a TX x TY tile of the x-y plane and marches through
// a chunk of z planes.  Cells are handled as pairs (i, i+1) with i odd: the
// two cells of a pair always have opposite colours, so each pair holds
// exactly one red cell.  LDS stores x_in by pair slot (first / second cell,
// each a dense array over pairs) and the new red values as one value per
// pair, so consecutive lanes touch consecutive LDS words; each lane picks
// its slot by address, not by branch.  LDS holds four x_in planes with a
// two-cell ring and three red planes with a one-cell ring.  Per plane m (two
// barriers):
//   red(m) on tile+ring | barrier | black(m-1) on the tile and store of
//   plane m-1; stage plane m+2 from registers into LDS | barrier
// Global loads of x (plane m+4) and b (red plane m+2, black plane m+1) are
// issued two plane steps before use into double-buffered registers (the
// plane loop is unrolled by two so buffer indices are compile-time).
// Blocks whose tile plus ring lies inside the domain (most of them) run a
// specialisation without bounds guards.  Coefficients of a thread's fixed
// pairs are kept in registers.
//
// Blocks are numbered with the z-chunk fastest, so with 8 (or a multiple of
// 8) chunks each XCD (blocks are dealt round-robin over the 8 XCDs) owns
// whole z-slabs and the tile rings are re-read from its own L2.
//
// MODE 0: plain.  MODE 1: x_in = 0 (the zeroed coarse x of
// src/multigrid.f90:1937; nothing is read).
#pragma once
#include <type_traits>

#include "pepsmg_kernels.h"

namespace pepsmg {

constexpr int FUSED_THREADS = 256;

template <typename T>
struct Vec2;
template <>
struct Vec2<double> { using type = double2; };
template <>
struct Vec2<float> { using type = float2; };

struct FusedGrid {
  int gx, gy, nch, kchunk;
};

// omega * (b - temp) / c0 + (1 - omega) * xold, the reference's SOR update
// (src/rbgs_poisson_matrix.f90:250).  FP64 divides exactly (bit-identical
// results); the FP32 inner solver of the mixed-precision path uses the
// hardware reciprocal-based divide.
__device__ __forceinline__ double sor_update(double omega, double bv, double temp, double c0, double xold) {
  return omega * (bv - temp) / c0 + (1.0 - omega) * xold;
}
__device__ __forceinline__ float sor_update(float omega, float bv, float temp, float c0, float xold) {
  return __fdividef(omega * (bv - temp), c0) + (1.0f - omega) * xold;
}

template <typename T, int TX, int TY>
struct FusedShape {
  static constexpr int NT = FUSED_THREADS;
  static constexpr int RP = TX / 2 + 2;            // pairs per row: x = i0-2 .. i0+TX+1
  static constexpr int RY = TY + 4;                // rows: y = j0-2 .. j0+TY+1
  static constexpr int QY = TY + 2;                // red rows: y = j0-1 .. j0+TY
  static constexpr int TP = TX / 2;                // tile pairs per row (pairs 1..TP)
  static constexpr int NXL = (RY * RP + NT - 1) / NT;
  static constexpr int NRED = (QY * RP + NT - 1) / NT;
  static constexpr int NBLK = (TY * TP + NT - 1) / NT;
  static constexpr int PLANE = RY * RP;
};

// IN: the tile and its two-cell ring lie inside the domain in x and y.
template <typename T, int TX, int TY, int MODE, int PF, bool IN>
__device__ __forceinline__ void rb_fused_body(T* __restrict__ xo, const T* __restrict__ xi, const T* __restrict__ b,
                                              const Lvl<T>& L, T omega, int i0, int j0, int k0, int k1,
                                              T* __restrict__ sa, T* __restrict__ sb, T* __restrict__ sr) {
  using S = FusedShape<T, TX, TY>;
  constexpr int NT = S::NT, RP = S::RP, RY = S::RY, QY = S::QY, TP = S::TP;
  constexpr int NXL = S::NXL, NRED = S::NRED, NBLK = S::NBLK, PLANE = S::PLANE;
  using V2 = typename Vec2<T>::type;
  const int t = threadIdx.x;

  // ---- red pairs of this thread: ring row qy, pair p ----
  int rq_y[NRED], rq_p[NRED];
  bool rq_ina[NRED], rq_inb[NRED];
  T rc_wa[NRED], rc_ea[NRED], rc_wb[NRED], rc_eb[NRED], rc_wy[NRED], rc_ey[NRED];
#pragma unroll
  for (int q = 0; q < NRED; q++) {
    int e = t + q * NT;
    int qy = e / RP, p = e - qy * RP;
    rq_y[q] = qy;
    rq_p[q] = p;
    int gj = j0 - 1 + qy, ia = i0 - 2 + 2 * p;
    bool row = e < QY * RP && (IN || (gj >= 1 && gj <= L.ny));
    // the first cell is used only from column i0-1 on, the second up to i0+TX
    rq_ina[q] = row && p > 0 && (IN || (ia >= 1 && ia <= L.nx));
    rq_inb[q] = row && p < RP - 1 && (IN || (ia + 1 >= 1 && ia + 1 <= L.nx));
    rc_wa[q] = rq_ina[q] ? L.cwx[ia] : T(0);
    rc_ea[q] = rq_ina[q] ? L.cex[ia] : T(0);
    rc_wb[q] = rq_inb[q] ? L.cwx[ia + 1] : T(0);
    rc_eb[q] = rq_inb[q] ? L.cex[ia + 1] : T(0);
    rc_wy[q] = row ? L.cwy[gj] : T(0);
    rc_ey[q] = row ? L.cey[gj] : T(0);
  }
  // ---- black pairs of this thread: tile row ty, pair p (1..TP) ----
  int bq_y[NBLK], bq_p[NBLK];
  bool bq_on[NBLK], bq_inb[NBLK];
  T bc_wa[NBLK], bc_ea[NBLK], bc_wb[NBLK], bc_eb[NBLK], bc_wy[NBLK], bc_ey[NBLK];
#pragma unroll
  for (int q = 0; q < NBLK; q++) {
    int e = t + q * NT;
    int ty = e / TP, p = 1 + (e - ty * TP);
    bq_y[q] = ty;
    bq_p[q] = p;
    int gj = j0 + ty, ia = i0 - 2 + 2 * p;
    bq_on[q] = e < TY * TP && (IN || (gj <= L.ny && ia <= L.nx));
    bq_inb[q] = bq_on[q] && (IN || ia + 1 <= L.nx);
    bc_wa[q] = bq_on[q] ? L.cwx[ia] : T(0);
    bc_ea[q] = bq_on[q] ? L.cex[ia] : T(0);
    bc_wb[q] = bq_inb[q] ? L.cwx[ia + 1] : T(0);
    bc_eb[q] = bq_inb[q] ? L.cex[ia + 1] : T(0);
    bc_wy[q] = bq_on[q] ? L.cwy[gj] : T(0);
    bc_ey[q] = bq_on[q] ? L.cey[gj] : T(0);
  }

  // ---- x_in plane loads (pairs); PF buffers (indexed by plane parity when PF = 2) ----
  V2 xp[PF][NXL];
  // Loads are issued unconditionally from clamped in-bounds addresses; the
  // values that lie outside the domain are zeroed only when they are used
  // (stage_x) or never read (b of cells outside the domain).  With no branch
  // or select right after a load the compiler waits on exactly the loads a
  // plane needs (vmcnt(n) with n > 0) instead of draining the prefetches.
  auto issue_x = [&](V2* dst, int m) {
    const int mc = min(max(m, 1), L.nz);
#pragma unroll
    for (int q = 0; q < NXL; q++) {
      int e = t + q * NT;
      int ec = min(e, RY * RP - 1);
      int ry = ec / RP, p = ec - ry * RP;
      int gi = i0 - 2 + 2 * p, gj = j0 - 2 + ry;    // gi is odd
      // an odd gi <= nx keeps the 2-element load inside the padded row
      int gic = IN ? gi : ((gi >= 1 && gi <= L.nx) ? gi : 1);
      int gjc = IN ? gj : min(max(gj, 1), L.ny);
      dst[q] = *reinterpret_cast<const V2*>(xi + lidx(L, gic, gjc, mc));
    }
  };
  auto stage_x = [&](const V2* src, int m) {
    const bool vz = m >= 1 && m <= L.nz;
#pragma unroll
    for (int q = 0; q < NXL; q++) {
      int e = t + q * NT;
      if (e < RY * RP) {
        int ry = e / RP, p = e - ry * RP;
        int gi = i0 - 2 + 2 * p, gj = j0 - 2 + ry;
        bool vrow = vz && (IN || (gj >= 1 && gj <= L.ny));
        bool va = vrow && (IN || (gi >= 1 && gi <= L.nx));
        bool vb = vrow && (IN || (gi + 1 >= 1 && gi + 1 <= L.nx));
        sa[(m & 3) * PLANE + e] = va ? src[q].x : T(0);
        sb[(m & 3) * PLANE + e] = vb ? src[q].y : T(0);
      }
    }
  };

  // red cell of pair (gj, m): the first cell (i odd) iff gj + m is even
  T br[PF][NRED], bk[PF][NBLK];
  // b of the red (black) cell of each pair; cells outside the domain load a
  // clamped address and their value is never used
  auto issue_bred = [&](T* dst, int m) {
    const int mc = min(max(m, 1), L.nz);
#pragma unroll
    for (int q = 0; q < NRED; q++) {
      int gj = j0 - 1 + rq_y[q];
      bool first = ((gj + m) & 1) == 0;
      bool in = first ? rq_ina[q] : rq_inb[q];
      int gi = i0 - 2 + 2 * rq_p[q] + (first ? 0 : 1);
      int gic = in ? gi : 1, gjc = in ? gj : 1;
      dst[q] = b[lidx(L, gic, gjc, mc)];
    }
  };
  auto issue_bblk = [&](T* dst, int k) {
    const int kc = min(max(k, 1), L.nz);
#pragma unroll
    for (int q = 0; q < NBLK; q++) {
      int gj = j0 + bq_y[q];
      bool bfirst = ((gj + k) & 1) == 1;     // black is the first cell iff gj + k odd
      bool in = bfirst ? bq_on[q] : bq_inb[q];
      int gi = i0 - 2 + 2 * bq_p[q] + (bfirst ? 0 : 1);
      int gic = in ? gi : 1, gjc = in ? gj : 1;
      dst[q] = b[lidx(L, gic, gjc, kc)];
    }
  };

  auto red_plane = [&](const T* bv, int m) {
    const int s0 = m & 3, sm = (m - 1) & 3, sp = (m + 1) & 3;
    T* r = sr + ((m + 3) % 3) * (QY * RP);
    bool zin = (m >= 1 && m <= L.nz);
    T c5 = T(0), c6 = T(0);
    if (zin) {
      c5 = ldc(L.cwz, m);
      c6 = ldc(L.cez, m);
    }
#pragma unroll
    for (int q = 0; q < NRED; q++) {
      int qy = rq_y[q], p = rq_p[q];
      if (qy >= QY) continue;
      int ry = qy + 1;                       // x_in row of this ring row
      int gj = j0 - 1 + qy;
      int first = ((gj + m) & 1) == 0;       // red cell is the first of its pair
      T v = T(0);
      if (zin && (first ? rq_ina[q] : rq_inb[q])) {
        T c1 = first ? rc_wa[q] : rc_wb[q], c2 = first ? rc_ea[q] : rc_eb[q];
        T c3 = rc_wy[q], c4 = rc_ey[q];
        T c0 = diag(c1, c2, c3, c4, c5, c6);
        // own slot SS (same column in other rows/planes), other slot OO
        // (x neighbours: pairs p-1,p if first, else p,p+1)
        const T* SS = first ? sa : sb;
        const T* OO = first ? sb : sa;
        int row = ry * RP + p;
        int ow = s0 * PLANE + row - first;
        T xw = OO[ow], xe = OO[ow + 1];
        T xs = SS[s0 * PLANE + row - RP], xn = SS[s0 * PLANE + row + RP];
        T xd = SS[sm * PLANE + row], xu = SS[sp * PLANE + row];
        T xc = SS[s0 * PLANE + row];
        T temp = c1 * xw + c2 * xe + c3 * xs + c4 * xn + c5 * xd + c6 * xu;
        v = sor_update(omega, bv[q], temp, c0, xc);
      }
      r[qy * RP + p] = v;
    }
  };

  auto black_plane = [&](const T* bv, int k) {
    const T* r0 = sr + (k % 3) * (QY * RP);
    const T* rm = sr + ((k + 2) % 3) * (QY * RP);
    const T* rp = sr + ((k + 1) % 3) * (QY * RP);
    const int s0 = k & 3;
    T c5 = ldc(L.cwz, k), c6 = ldc(L.cez, k);
#pragma unroll
    for (int q = 0; q < NBLK; q++) {
      if (!bq_on[q]) continue;
      int ty = bq_y[q], p = bq_p[q];
      int gj = j0 + ty;
      int ry = ty + 1;                       // red row of tile row ty
      int bfirst = ((gj + k) & 1) == 1;      // black cell is the first of its pair
      bool in = bfirst ? true : bq_inb[q];
      int row = ry * RP + p;
      T vb = T(0);
      if (in) {
        T c1 = bfirst ? bc_wa[q] : bc_wb[q], c2 = bfirst ? bc_ea[q] : bc_eb[q];
        T c3 = bc_wy[q], c4 = bc_ey[q];
        T c0 = diag(c1, c2, c3, c4, c5, c6);
        int ow = row - bfirst;               // red x neighbours: pairs p-1,p or p,p+1
        T xw = r0[ow], xe = r0[ow + 1];
        T temp = c1 * xw + c2 * xe + c3 * r0[row - RP] + c4 * r0[row + RP] + c5 * rm[row] + c6 * rp[row];
        const T* SS = bfirst ? sa : sb;
        T xc = SS[s0 * PLANE + (ty + 2) * RP + p];
        vb = sor_update(omega, bv[q], temp, c0, xc);
      }
      T vr = r0[row];
      T lo = bfirst ? vb : vr, hi = bfirst ? vr : vb;
      long q0 = lidx(L, i0 - 2 + 2 * p, gj, k);
      if (bq_inb[q]) {
        V2 w;
        w.x = lo;
        w.y = hi;
        *reinterpret_cast<V2*>(xo + q0) = w;
      } else {
        xo[q0] = lo;
      }
    }
  };

  // prologue: planes k0-2 .. k0 staged; k0+1 .. k0+PF in flight
  if (MODE == 1) {
    for (int e = t; e < 4 * PLANE; e += NT) {
      sa[e] = T(0);
      sb[e] = T(0);
    }
  } else {
    issue_x(xp[0], k0 - 2);
    stage_x(xp[0], k0 - 2);
    issue_x(xp[0], k0 - 1);
    stage_x(xp[0], k0 - 1);
    issue_x(xp[0], k0);
    stage_x(xp[0], k0);
    if (PF == 2) {
      issue_x(xp[1 % PF], k0 + 1);
      issue_x(xp[0], k0 + 2);
    } else {
      issue_x(xp[0], k0 + 1);
    }
  }
  if (PF == 2) {
    issue_bred(br[1 % PF], k0 - 1);
    issue_bred(br[0], k0);
    issue_bblk(bk[0], k0);
  } else {
    issue_bred(br[0], k0 - 1);
  }

  // One plane step.  Loads for plane q are issued PF steps before plane q is
  // used; with PF = 2 the buffers alternate by plane parity
  // (P = (m - k0) & 1, compile-time through the two-step unrolled loop).
  auto step = [&](int m, auto Pc) {
    constexpr int P = PF == 2 ? decltype(Pc)::value : 0;
    constexpr int PB = PF == 2 ? 1 - P : 0;
    lds_barrier();
    red_plane(br[P], m);                     // needs b(m), x planes m-1..m+1
    issue_bred(br[P], m + PF);
    lds_barrier();
    if (m - 1 >= k0) black_plane(bk[PB], m - 1);
    if (m >= k0) issue_bblk(bk[PB], m - 1 + PF);
    if (MODE != 1) {
      stage_x(xp[P], m + 2);
      issue_x(xp[P], m + 2 + PF);
    }
  };
  for (int m = k0 - 1; m <= k1; m += 2) {
    step(m, std::integral_constant<int, 1>());
    if (m + 1 <= k1) step(m + 1, std::integral_constant<int, 0>());
  }
}

// One block of the fused iteration; bid numbers the blocks of one grid.
template <typename T, int TX, int TY, int MODE, int PF>
__device__ __forceinline__ void rb_fused_block(int bid, T* __restrict__ xo, const T* __restrict__ xi,
                                               const T* __restrict__ b, const Lvl<T>& L, T omega,
                                               const FusedGrid& G) {
  using S = FusedShape<T, TX, TY>;
  __shared__ T sa[4 * S::PLANE];   // x_in, first cell of each pair (i = i0-2+2p, odd)
  __shared__ T sb[4 * S::PLANE];   // x_in, second cell (i+1)
  __shared__ T sr[3 * S::QY * S::RP];   // new red value of each pair

  const int bz = bid % G.nch;
  const int rxy = bid / G.nch;
  const int i0 = 1 + (rxy % G.gx) * TX;
  const int j0 = 1 + (rxy / G.gx) * TY;
  const int k0 = 1 + bz * G.kchunk;
  const int k1 = min(k0 + G.kchunk, L.nz + 1);   // black planes k0 .. k1-1
  const bool in = i0 >= 3 && i0 + TX + 1 <= L.nx && j0 >= 3 && j0 + TY + 1 <= L.ny;
  if (in)
    rb_fused_body<T, TX, TY, MODE, PF, true>(xo, xi, b, L, omega, i0, j0, k0, k1, sa, sb, sr);
  else
    rb_fused_body<T, TX, TY, MODE, PF, false>(xo, xi, b, L, omega, i0, j0, k0, k1, sa, sb, sr);
}

template <typename T, int TX, int TY, int MODE, int PF>
__global__ void __launch_bounds__(FUSED_THREADS)
k_rb_fused(T* __restrict__ xo, const T* __restrict__ xi, const T* __restrict__ b, Lvl<T> L, T omega, FusedGrid G) {
  rb_fused_block<T, TX, TY, MODE, PF>(blockIdx.x, xo, xi, b, L, omega, G);
}

}  // namespace pepsmg
