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
// two-cell ring and four red planes with a one-cell ring.  Per plane m (two
// barriers):
//   red(m) on tile+ring | barrier | black(m-1) on the tile and store of
//   plane m-1; stage plane m+2 from registers into LDS | barrier
// Global loads of x (plane m+2+PF) and b (red plane m+PF, black plane
// m-1+PF) are issued PF plane steps before use (PF = 1 or 2, registers
// double-buffered for 2).  Per-thread data that depends on the plane only
// through its parity (which cell of a pair is red, coefficients, partial
// diagonal sums, LDS and b offsets) is precomputed for both parities.
// Workgroups map to (tile, chunk) with the chunk fastest or, for tiles of
// 2048 cells or more, XCD-aware (fused_xcd_bid).
//
// MODE 0: plain.  MODE 1: x_in = 0 (the zeroed coarse x of
// src/multigrid.f90:1937; x is not read): the LDS planes and loads that
// stage x_in stage b instead, so b arrives as whole pairs once per plane
// (the other modes load it per colour into registers).  MODE 2: plain, and each block
// also writes the partial sum of (b - A x_in)^2 over its tile cells (the
// convergence norm of src/multigrid.f90:1970-1971 for the x the iteration
// starts from): the residual of plane m is evaluated in the red phase of
// step m, when x_in planes m-1 .. m+1 are in LDS, with mv_mul's expression
// (src/poisson_matrix_operator.f90:150-156).
// FP64 only, the ends of a solve on the caller's padded (0:n+1)^3 arrays
// (layout U: pitches nx+2 and (nx+2)(ny+2), base at cell (1,1,1); their
// pairs (i odd, i+1) are only 8-byte aligned, accessed as unaligned 16-byte
// vectors):
// MODE 3: MODE 2 with x_in and b read from the caller's arrays, and each
//   tile's b stored into the solver's layout (ext), so the solve needs no
//   copy-in pass.
// MODE 4: MODE 0 that also stores x_out into the caller's array (ext), for
//   the iteration expected to end the solve (no copy-out pass).
#pragma once
#include <type_traits>

#include "pepsmg_kernels.h"

namespace pepsmg {

constexpr int FUSED_THREADS = 256;
// threads of a TX x TY tile: two black pairs per thread (at least 256
// threads), so taller or wider tiles keep the per-thread registers of the
// 64 x 16 and 128 x 8 ones
constexpr int fused_threads(int tx, int ty) { return tx * ty / 4 > FUSED_THREADS ? tx * ty / 4 : FUSED_THREADS; }

template <typename T>
struct Vec2;
template <>
struct Vec2<double> { using type = double2; };
template <>
struct Vec2<float> { using type = float2; };

// two adjacent elements (16-byte / 8-byte aligned) loaded non-temporally
template <typename T>
__device__ __forceinline__ typename Vec2<T>::type ld_stream_v2(const T* p) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  v2 r = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  typename Vec2<T>::type o;
  o.x = r.x;
  o.y = r.y;
  return o;
}

// Two adjacent elements at an address aligned to one element only (the
// pairs (i odd, i+1) of the caller's padded rows): one 16-byte (FP64) access,
// which gfx950 performs unaligned, instead of two element accesses that each
// cover every other element of a cache line.
template <typename T>
__device__ __forceinline__ typename Vec2<T>::type ld_pair_unal(const T* p) {
  typedef T v2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
  const v2 r = *reinterpret_cast<const v2*>(p);
  typename Vec2<T>::type o;
  o.x = r.x;
  o.y = r.y;
  return o;
}
template <typename T>
__device__ __forceinline__ void st_pair_unal(T* p, T a, T b) {
  typedef T v2 __attribute__((ext_vector_type(2), aligned(sizeof(T))));
  v2 v = {a, b};
  *reinterpret_cast<v2*>(p) = v;
}

struct FusedGrid {
  int gx, gy, nch, kchunk;
  int order = 0;   // workgroup order (fused_xcd_bid)
};

// omega * (b - temp) / c0 + (1 - omega) * xold, the reference's SOR update
// (src/rbgs_poisson_matrix.f90:250).  FP64 divides exactly (bit-identical
// results); the FP32 inner solver of the mixed-precision path multiplies by
// the hardware reciprocal (1 ulp), which leaves its convergence unchanged.
__device__ __forceinline__ double sor_update(double omega, double bv, double temp, double c0, double xold) {
  return omega * (bv - temp) / c0 + (1.0 - omega) * xold;
}
__device__ __forceinline__ float sor_update(float omega, float bv, float temp, float c0, float xold) {
  return omega * (bv - temp) * __builtin_amdgcn_rcpf(c0) + (1.0f - omega) * xold;
}

// Modes 0 and 4 on tiles of 2048 cells or more (one or two blocks per CU
// either way) load b by whole pairs along with the red values and pass the
// black cells' b to the black phase through LDS, instead of a load per
// colour (FP64 512^3 plane iterations 1.4 % faster, mode 4 6 %); the norm
// modes 2 and 3 ran 18 % slower that way and keep the loads per colour, and
// mode 1 stages b in the x planes instead.
constexpr bool rb_pair_b(int mode, int tx, int ty) { return (mode == 0 || mode == 4) && tx * ty >= 2048; }

template <typename T, int TX, int TY>
struct FusedShape {
  static constexpr int NT = fused_threads(TX, TY);
  static constexpr int RP = TX / 2 + 2;            // pairs per row: x = i0-2 .. i0+TX+1
  static constexpr int RY = TY + 4;                // rows: y = j0-2 .. j0+TY+1
  static constexpr int QY = TY + 2;                // red rows: y = j0-1 .. j0+TY
  static constexpr int TP = TX / 2;                // tile pairs per row (pairs 1..TP)
  static constexpr int NXL = (RY * RP + NT - 1) / NT;
  static constexpr int NRED = (QY * RP + NT - 1) / NT;
  static constexpr int NBLK = (TY * TP + NT - 1) / NT;
  static constexpr int PLANE = RY * RP;
};

// Everything that depends on a thread's fixed pairs is computed once per
// block, for both plane parities z: which cell of a pair is red (black) on
// planes of parity z, its x coefficients, the partial diagonal sum
// ((c1 + c2) + c3) + c4 (c0 = -((s4 + c5) + c6) is the reference's
// left-to-right sum), its LDS offsets and the offset of its b value.  The
// chunk start k0 is odd (even chunk length), so each plane step of the
// two-step unrolled loop knows its plane parity at compile time.
// DOWN (mode 0, PF 1): the march runs from the top of the chunk down; every
// cell gets the same value (the colours' order fixes it, not the planes').
template <typename T, int TX, int TY, int MODE, int PF, bool DOWN = false>
__device__ __forceinline__ double rb_fused_body(T* __restrict__ xo, const T* __restrict__ xi,
                                                const T* __restrict__ b, const Lvl<T>& L, T omega, int i0, int j0,
                                                int k0, int k1, T* __restrict__ sx, T* __restrict__ sr,
                                                T* __restrict__ sbk, const Lvl<T>& U, T* __restrict__ ext) {
  constexpr bool NORM = MODE == 2 || MODE == 3;   // partial sums of (b - A x_in)^2
  constexpr bool CALLER_IN = MODE == 3;           // x_in and b in the caller's layout U
  constexpr bool CALLER_OUT = MODE == 4;          // x_out also into the caller's layout U
  constexpr bool ZB = MODE == 1;                  // x_in = 0; the staged planes hold b
  constexpr bool PB = rb_pair_b(MODE, TX, TY);     // b loaded by whole pairs with the red values
  using S = FusedShape<T, TX, TY>;
  constexpr int NT = S::NT, RP = S::RP, QY = S::QY, TP = S::TP;
  constexpr int NXL = S::NXL, NRED = S::NRED, NBLK = S::NBLK, PLANE = S::PLANE;
  constexpr int SB = 4 * PLANE;           // second-cell slots follow the four first-cell slots
  constexpr int QP = QY * RP;             // one red plane
  using V2 = typename Vec2<T>::type;
  const int t = threadIdx.x;
  const long pxy = L.pxy;
  // pitches of the inputs x_in and b
  const int ipx = CALLER_IN ? (int)U.px : (int)L.px;
  const long ipxy = CALLER_IN ? U.pxy : L.pxy;

  // ---- red pairs: ring row qy (y = j0-1+qy), pair p (x = i0-2+2p, +1) ----
  bool r_in[NRED][2];
  T r_c1[NRED][2], r_c2[NRED][2], r_s4[NRED][2], r_c3[NRED], r_c4[NRED];
  int r_ss[NRED][2], r_oo[NRED][2], r_bo[NRED][2], r_q[NRED];
  // PB: offset of the pair's b (in bounds), the tile slot of its black cell
  // (-1 on the ring) and whether the red cell is the pair's first
  int r_po[PB ? NRED : 1], r_bt[PB ? NRED : 1];
  bool r_first[PB ? NRED : 1][2];
#pragma unroll
  for (int q = 0; q < NRED; q++) {
    int e = min(t + q * NT, QP - 1);
    int qy = e / RP, p = e - qy * RP;
    int gj = j0 - 1 + qy, ia = i0 - 2 + 2 * p;
    bool row = t + q * NT < QP && gj >= 1 && gj <= L.ny;
    if (PB) {
      // an odd column <= nx keeps the 2-element load inside the padded row
      r_po[q] = (row && ia >= 1 && ia <= L.nx) ? (gj - 1) * ipx + (ia - 1) : 0;
      r_bt[q] = (t + q * NT < QP && qy >= 1 && qy <= TY && p >= 1 && p <= TP) ? (qy - 1) * TP + (p - 1) : -1;
    }
    // the first cell is used only from column i0-1 on, the second up to i0+TX
    bool ina = row && p > 0 && ia >= 1 && ia <= L.nx;
    bool inb = row && p < RP - 1 && ia + 1 >= 1 && ia + 1 <= L.nx;
    T c3 = row ? L.cwy[gj] : T(0), c4 = row ? L.cey[gj] : T(0);
    r_c3[q] = c3;
    r_c4[q] = c4;
    r_q[q] = e;
    int rrow = (qy + 1) * RP + p;           // x_in row of this ring row
#pragma unroll
    for (int z = 0; z < 2; z++) {
      bool first = ((gj + z) & 1) == 0;     // red is the first cell iff gj + k even
      if (PB) r_first[q][z] = first;
      bool in = first ? ina : inb;
      int ic = first ? ia : ia + 1;
      T c1 = in ? L.cwx[ic] : T(0), c2 = in ? L.cex[ic] : T(0);
      r_in[q][z] = in;
      r_c1[q][z] = c1;
      r_c2[q][z] = c2;
      r_s4[q][z] = c1 + c2 + c3 + c4;
      r_ss[q][z] = (first ? 0 : SB) + rrow;
      r_oo[q][z] = (first ? SB : 0) + rrow - (first ? 1 : 0);   // west neighbour; east is +1
      r_bo[q][z] = in ? (gj - 1) * ipx + (ic - 1) : 0;
    }
  }
  // ---- black pairs: tile row ty (y = j0+ty), pair p = 1..TP ----
  bool b_on[NBLK], b_odd[NBLK], b_in[NBLK][2];
  T b_c1[NBLK][2], b_c2[NBLK][2], b_s4[NBLK][2], b_c3[NBLK], b_c4[NBLK];
  int b_row[NBLK], b_ow[NBLK][2], b_xc[NBLK][2], b_bo[NBLK][2], b_so[NBLK];
  int b_xw[NORM ? NBLK : 1][2];             // MODE 2, 3: x_in index of the west neighbour of b_xc
  constexpr bool USI = CALLER_IN || CALLER_OUT;
  int b_si[USI ? NBLK : 1];                 // MODE 3, 4: the pair's first cell in the caller's layout
  bool b_two[USI ? NBLK : 1];               // MODE 3, 4: the pair's second cell lies inside
#pragma unroll
  for (int q = 0; q < NBLK; q++) {
    int e = min(t + q * NT, TY * TP - 1);
    int ty = e / TP, p = 1 + (e - ty * TP);
    int gj = j0 + ty, ia = i0 - 2 + 2 * p;
    bool on = t + q * NT < TY * TP && gj <= L.ny && ia <= L.nx;
    bool two = on && ia + 1 <= L.nx;
    b_on[q] = on;
    b_odd[q] = (gj & 1) == 1;
    T c3 = on ? L.cwy[gj] : T(0), c4 = on ? L.cey[gj] : T(0);
    b_c3[q] = c3;
    b_c4[q] = c4;
    b_row[q] = (ty + 1) * RP + p;           // red-ring row of tile row ty
    b_so[q] = on ? (gj - 1) * (int)L.px + (ia - 1) : 0;
    if (USI) {
      b_si[q] = on ? (gj - 1) * (int)U.px + (ia - 1) : 0;
      b_two[q] = two;
    }
#pragma unroll
    for (int z = 0; z < 2; z++) {
      bool bfirst = ((gj + z) & 1) == 1;    // black is the first cell iff gj + k odd
      bool in = bfirst ? on : two;
      int ic = bfirst ? ia : ia + 1;
      T c1 = in ? L.cwx[ic] : T(0), c2 = in ? L.cex[ic] : T(0);
      b_in[q][z] = in;
      b_c1[q][z] = c1;
      b_c2[q][z] = c2;
      b_s4[q][z] = c1 + c2 + c3 + c4;
      b_ow[q][z] = b_row[q] - (bfirst ? 1 : 0);   // red x neighbours: pairs p-1,p or p,p+1
      b_xc[q][z] = (bfirst ? 0 : SB) + (ty + 2) * RP + p;
      b_bo[q][z] = in ? (gj - 1) * ipx + (ic - 1) : 0;
      if (NORM) b_xw[q][z] = bfirst ? SB + (ty + 2) * RP + p - 1 : (ty + 2) * RP + p;
    }
  }
  // ---- x_in staging: pair e of a plane (row ry, pair p) ----
  int x_off[NXL];
  unsigned x_ok = 0;                        // bit 2q: first cell of item q inside, 2q+1: second
#pragma unroll
  for (int q = 0; q < NXL; q++) {
    int e = min(t + q * NT, PLANE - 1);
    int ry = e / RP, p = e - ry * RP;
    int gi = i0 - 2 + 2 * p, gj = j0 - 2 + ry;    // gi is odd
    bool vrow = gj >= 1 && gj <= L.ny;
    if (vrow && gi >= 1 && gi <= L.nx) x_ok |= 1u << (2 * q);
    if (vrow && gi + 1 >= 1 && gi + 1 <= L.nx) x_ok |= 2u << (2 * q);
    // an odd gi <= nx keeps the 2-element load inside the padded row
    int gic = (gi >= 1 && gi <= L.nx) ? gi : 1;
    int gjc = min(max(gj, 1), L.ny);
    x_off[q] = (gjc - 1) * ipx + (gic - 1);
  }

  // Loads are issued unconditionally (clamped plane, precomputed in-bounds
  // offsets); values outside the domain are zeroed when staged or never
  // read.  With no branch or select next to a load the compiler waits only
  // for the loads a plane needs instead of draining the prefetches.
  V2 xp[PF][NXL];
  T br[PB ? 1 : PF][NRED], bk[PB ? 1 : PF][NBLK];
  V2 brp[PB ? PF : 1][NRED];
  auto issue_x = [&](V2* dst, int m) {
    const T* base = (ZB ? b : xi) + (long)(min(max(m, 1), L.nz) - 1) * ipxy;
#pragma unroll
    for (int q = 0; q < NXL; q++) {
      if (CALLER_IN) {
        dst[q] = ld_pair_unal(base + x_off[q]);
      } else {
        dst[q] = *reinterpret_cast<const V2*>(base + x_off[q]);
      }
    }
  };
  auto stage_x = [&](const V2* src, int m) {
    const unsigned ok = (m >= 1 && m <= L.nz) ? x_ok : 0u;
    T* d = sx + (m & 3) * PLANE;
#pragma unroll
    for (int q = 0; q < NXL; q++) {
      int e = t + q * NT;
      if (e < PLANE) {
        d[e] = ((ok >> (2 * q)) & 1u) ? src[q].x : T(0);
        d[SB + e] = ((ok >> (2 * q)) & 2u) ? src[q].y : T(0);
      }
    }
  };
  // b of the red cells of plane m into buffer slot (PB: whole pairs)
  auto issue_bred = [&](auto slot_c, int m, auto Zc) {
    if (ZB) return;
    constexpr int slot = decltype(slot_c)::value;
    constexpr int Z = decltype(Zc)::value;
    const T* base = b + (long)(min(max(m, 1), L.nz) - 1) * ipxy;
#pragma unroll
    for (int q = 0; q < NRED; q++) {
      if (PB) {
        if (CALLER_IN) brp[slot][q] = ld_pair_unal(base + r_po[q]);
        else brp[slot][q] = *reinterpret_cast<const V2*>(base + r_po[q]);
      } else {
        br[slot][q] = base[r_bo[q][Z]];
      }
    }
  };
  auto issue_bblk = [&](T* dst, int k, auto Zc) {
    if (ZB || PB) return;
    constexpr int Z = decltype(Zc)::value;
    const T* base = b + (long)(min(max(k, 1), L.nz) - 1) * ipxy;
#pragma unroll
    for (int q = 0; q < NBLK; q++) dst[q] = base[b_bo[q][Z]];
  };

  // new red values of plane m (parity Z) on the tile and its ring
  // PB: the black cells' b of the tile pairs go to sbk (plane parity slot)
  // for black_plane one step later
  auto red_plane = [&](auto slot_c, int m, auto Zc) {
    constexpr int slot = decltype(slot_c)::value;
    constexpr int Z = decltype(Zc)::value;
    const T* bv = br[PB ? 0 : slot];
    const T* X0 = sx + (m & 3) * PLANE;
    const T* Xm = sx + ((m - 1) & 3) * PLANE;
    const T* Xp = sx + ((m + 1) & 3) * PLANE;
    T* r = sr + (m & 3) * QP;
    const bool zin = m >= 1 && m <= L.nz;
    const int mc = min(max(m, 1), L.nz);
    const T c5 = ldc(L.cwz, mc), c6 = ldc(L.cez, mc);
#pragma unroll
    for (int q = 0; q < NRED; q++) {
      if (q * NT >= QP) continue;
      T v = T(0);
      if (zin && r_in[q][Z]) {
        const int ss = r_ss[q][Z], oo = r_oo[q][Z];
        T c1 = r_c1[q][Z], c2 = r_c2[q][Z], c3 = r_c3[q], c4 = r_c4[q];
        T c0 = -(r_s4[q][Z] + c5 + c6);
        if (ZB) {   // black neighbours and the old value are 0 (their products sum to +0)
          v = sor_update(omega, X0[ss], T(0), c0, T(0));
        } else {
          T temp = c1 * X0[oo] + c2 * X0[oo + 1] + c3 * X0[ss - RP] + c4 * X0[ss + RP] + c5 * Xm[ss] + c6 * Xp[ss];
          T bred = bv[q];
          if (PB) {
            const T p0 = brp[slot][q].x, p1 = brp[slot][q].y;
            bred = r_first[q][Z] ? p0 : p1;
          }
          v = sor_update(omega, bred, temp, c0, X0[ss]);
        }
      }
      if (t + q * NT < QP) r[r_q[q]] = v;
      if (PB && r_bt[q] >= 0) {
        const T p0 = brp[slot][q].x, p1 = brp[slot][q].y;
        sbk[(m & 1) * TY * TP + r_bt[q]] = r_first[q][Z] ? p1 : p0;
      }
    }
  };

  // black values of plane k (parity Z) on the tile, stored with the red ones
  auto black_plane = [&](const T* bv, int k, auto Zc) {
    constexpr int Z = decltype(Zc)::value;
    const T* sb = sbk + (k & 1) * TY * TP;   // PB: b of the tile's black cells of plane k
    const T* r0 = sr + (k & 3) * QP;
    const T* rm = sr + ((k - 1) & 3) * QP;
    const T* rp = sr + ((k + 1) & 3) * QP;
    const T* X0 = sx + (k & 3) * PLANE;
    const T c5 = ldc(L.cwz, k), c6 = ldc(L.cez, k);
    T* out = xo + (long)(k - 1) * pxy;
#pragma unroll
    for (int q = 0; q < NBLK; q++) {
      if (!b_on[q]) continue;
      const int row = b_row[q];
      T vb = T(0);
      if (b_in[q][Z]) {
        const int ow = b_ow[q][Z];
        T c1 = b_c1[q][Z], c2 = b_c2[q][Z], c3 = b_c3[q], c4 = b_c4[q];
        T c0 = -(b_s4[q][Z] + c5 + c6);
        T temp = c1 * r0[ow] + c2 * r0[ow + 1] + c3 * r0[row - RP] + c4 * r0[row + RP] + c5 * rm[row] + c6 * rp[row];
        if (ZB) vb = sor_update(omega, X0[b_xc[q][Z]], temp, c0, T(0));   // X0: b of plane k
        else vb = sor_update(omega, PB ? sb[min(t + q * NT, TY * TP - 1)] : bv[q], temp, c0, X0[b_xc[q][Z]]);
      }
      T vr = r0[row];
      const bool bf = Z ? !b_odd[q] : b_odd[q];   // black is the first cell iff gj + k odd
      // the pair is stored whole: a second cell beyond nx lies in the row's
      // padding (odd column <= nx, pitch a multiple of 16) and gets 0
      st_stream_pair(out + b_so[q], bf ? vb : vr, bf ? vr : vb);
      if (CALLER_OUT) {   // the caller's row holds the ghost cell nx+1 past the last pair's first cell
        T* uo = ext + (long)(k - 1) * U.pxy + b_si[q];
        if (b_two[q]) st_pair_unal(uo, bf ? vb : vr, bf ? vr : vb);
        else uo[0] = bf ? vb : vr;
      }
    }
  };

  // MODE 2: (b - A x_in)^2 of the tile cells of plane m (parity Z), both
  // cells of each pair: the black cell of parity Z is b_xc[q][Z], the red one
  // b_xc[q][1-Z] (the black cell of the other parity), each with its own
  // coefficients; b of the pair arrives as one aligned 2-element load.
  V2 bn[NORM ? NBLK : 1];
  double acc = 0.0;
  auto issue_bn = [&](int m) {
    if (!NORM) return;
    const T* base = b + (long)(min(max(m, 1), L.nz) - 1) * ipxy;
#pragma unroll
    for (int q = 0; q < NBLK; q++) {
      if (CALLER_IN) {
        bn[q] = ld_pair_unal(base + b_si[q]);
      } else {
        bn[q] = *reinterpret_cast<const V2*>(base + b_so[q]);
      }
    }
  };
  auto norm_plane = [&](int m, auto Zc) {
    if (!NORM) return;
    constexpr int Z = decltype(Zc)::value;
    const T* X0 = sx + (m & 3) * PLANE;
    const T* Xm = sx + ((m - 1) & 3) * PLANE;
    const T* Xp = sx + ((m + 1) & 3) * PLANE;
    const T c5 = ldc(L.cwz, m), c6 = ldc(L.cez, m);
#pragma unroll
    for (int q = 0; q < NBLK; q++) {
      if (!b_on[q]) continue;
      const bool bf = Z ? !b_odd[q] : b_odd[q];   // black is the first cell iff gj + k odd
#pragma unroll
      for (int w = 0; w < 2; w++) {
        const int zz = w == 0 ? Z : 1 - Z;          // w 0: black cell, 1: red cell
        if (!b_in[q][zz]) continue;
        const int c = b_xc[q][zz], ow = b_xw[q][zz];
        const T c1 = b_c1[q][zz], c2 = b_c2[q][zz], c3 = b_c3[q], c4 = b_c4[q];
        const T c0 = -(b_s4[q][zz] + c5 + c6);
        const T ax = c0 * X0[c] + c1 * X0[ow] + c2 * X0[ow + 1] + c3 * X0[c - RP] + c4 * X0[c + RP] + c5 * Xm[c] +
                     c6 * Xp[c];
        const bool first = (w == 0) == bf;
        const T r = (first ? bn[q].x : bn[q].y) - ax;
        acc += (double)r * (double)r;
      }
    }
    if (CALLER_IN) {   // the tile's b into the solver's layout (a second cell beyond nx gets 0)
      T* bo = ext + (long)(m - 1) * pxy;
#pragma unroll
      for (int q = 0; q < NBLK; q++)
        if (b_on[q]) st_stream_pair(bo + b_so[q], bn[q].x, b_two[q] ? bn[q].y : T(0));
    }
  };

  using Z0 = std::integral_constant<int, 0>;
  using Z1 = std::integral_constant<int, 1>;
  if constexpr (DOWN) {
    static_assert(MODE == 0 && PF == 1, "downward march: plain iterations, one plane of look-ahead");
    // Mirror of the upward march below: plane step m computes red(m) and
    // then black(m+1), stages x(m-2) and issues x(m-3).  It starts at the
    // odd plane mtop >= k1 (k1 is even only at the top of the domain, where
    // plane nz+2 gives zeros), so every step knows its plane parity at
    // compile time, and ends at k0-1.
    const int mtop = k1 | 1;
    {
      V2 pa[NXL], pb[NXL];
      issue_x(xp[0], mtop + 1);
      issue_x(pa, mtop);
      issue_x(pb, mtop - 1);
      stage_x(xp[0], mtop + 1);
      stage_x(pa, mtop);
      stage_x(pb, mtop - 1);
    }
    issue_x(xp[0], mtop - 2);
    issue_bred(std::integral_constant<int, 0>(), mtop, Z1());
    auto step_dn = [&](int m, auto Zc) {
      constexpr int ZM = decltype(Zc)::value;   // parity of m
      lds_barrier();
      red_plane(std::integral_constant<int, 0>(), m, Zc);   // b(m), x planes m-1..m+1
      issue_bred(std::integral_constant<int, 0>(), m - 1, std::integral_constant<int, 1 - ZM>());
      lds_barrier();
      if (m + 1 <= k1 - 1) black_plane(bk[0], m + 1, std::integral_constant<int, 1 - ZM>());
      if (m <= k1 - 1) issue_bblk(bk[0], m, Zc);
      stage_x(xp[0], m - 2);
      issue_x(xp[0], m - 3);
    };
    for (int m = mtop; m >= k0; m -= 2) {
      step_dn(m, Z1());
      step_dn(m - 1, Z0());
    }
    return 0.0;
  }
  // prologue: planes k0-2 .. k0 staged; k0+1 .. k0+PF in flight (k0 is
  // odd).  The three staged planes' loads are all issued before the first is
  // staged: one memory round trip instead of three (small levels are bound
  // by that latency).
  {
    V2 pa[NXL], pb[NXL];
    issue_x(xp[0], k0 - 2);
    issue_x(pa, k0 - 1);
    issue_x(pb, k0);
    stage_x(xp[0], k0 - 2);
    stage_x(pa, k0 - 1);
    stage_x(pb, k0);
  }
  if (PF == 2) {
    issue_x(xp[1 % PF], k0 + 1);
    issue_x(xp[0], k0 + 2);
  } else {
    issue_x(xp[0], k0 + 1);
  }
  if (PF == 2) {
    issue_bred(std::integral_constant<int, 1 % PF>(), k0 - 1, Z0());
    issue_bred(std::integral_constant<int, 0>(), k0, Z1());
    issue_bblk(bk[0], k0, Z1());
  } else {
    issue_bred(std::integral_constant<int, 0>(), k0 - 1, Z0());
  }

  // One plane step for plane m: red(m) | barrier | black(m-1), stage x(m+2).
  // P = (m - k0) & 1, so plane m has parity 1 - P.  Loads for a plane are
  // issued PF steps before it is used; with PF = 2 the buffers alternate by P.
  auto step = [&](int m, auto Pc) {
    constexpr int P = decltype(Pc)::value;
    constexpr int ZM = 1 - P;                       // parity of m
    constexpr int B = PF == 2 ? P : 0, BB = PF == 2 ? 1 - P : 0;
    lds_barrier();
    red_plane(std::integral_constant<int, B>(), m, std::integral_constant<int, ZM>());   // b(m), x planes m-1..m+1
    issue_bred(std::integral_constant<int, B>(), m + PF, std::integral_constant<int, (ZM + PF) & 1>());
    if (NORM) {
      if (m >= k0 && m < k1) norm_plane(m, std::integral_constant<int, ZM>());
      issue_bn(m + 1);
    }
    lds_barrier();
    if (m - 1 >= k0) black_plane(bk[PB ? 0 : BB], m - 1, std::integral_constant<int, 1 - ZM>());
    if (m >= k0) issue_bblk(bk[BB], m - 1 + PF, std::integral_constant<int, (1 - ZM + PF) & 1>());
    stage_x(xp[B], m + 2);
    issue_x(xp[B], m + 2 + PF);
  };
  for (int m = k0 - 1; m <= k1; m += 2) {
    step(m, std::integral_constant<int, 1>());
    if (m + 1 <= k1) step(m + 1, std::integral_constant<int, 0>());
  }
  return acc;
}

// One block of the fused iteration; bid numbers the blocks of one grid.
template <typename T, int TX, int TY, int MODE, int PF, bool DOWN = false>
__device__ __forceinline__ void rb_fused_block(int bid, T* __restrict__ xo, const T* __restrict__ xi,
                                               const T* __restrict__ b, const Lvl<T>& L, T omega,
                                               const FusedGrid& G, double* __restrict__ partial, const Lvl<T>& U,
                                               T* __restrict__ ext) {
  using S = FusedShape<T, TX, TY>;
  // x_in planes (4 slots): first cells of the pairs (i = i0-2+2p, odd),
  // then second cells (i+1); new red values (4 slots), one per pair
  __shared__ T sx[8 * S::PLANE];
  __shared__ T sr[4 * S::QY * S::RP];
  __shared__ T sbk[rb_pair_b(MODE, TX, TY) ? 2 * TY * S::TP : 1];   // two planes of black-cell b

  const int bz = bid % G.nch;
  const int rxy = bid / G.nch;
  const int i0 = 1 + (rxy % G.gx) * TX;
  const int j0 = 1 + (rxy / G.gx) * TY;
  const int k0 = 1 + bz * G.kchunk;              // odd: kchunk is even
  const int k1 = min(k0 + G.kchunk, L.nz + 1);   // black planes k0 .. k1-1
  const double acc =
      rb_fused_body<T, TX, TY, MODE, PF, DOWN>(xo, xi, b, L, omega, i0, j0, k0, k1, sx, sr, sbk, U, ext);
  if (MODE == 2 || MODE == 3) {
    __shared__ double sh[fused_threads(TX, TY) / 64];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[bid] = s;
  }
}

// Block number of workgroup w in a grid of nb blocks.  order 0: w itself
// (z-chunk fastest).  order 1, XCD-aware: the hardware deals workgroups
// round-robin over the 8 XCDs, so XCD w % 8 gets the w / 8-th block of a
// contiguous eighth of the order x-tiles fastest, then y-tiles, then
// z-chunks: each XCD owns full-width bands of tiles of one slab and re-reads
// the column rings beside a tile from its own L2.  The block number keeps
// the chunk fastest either way (the partial-sum slot of a tile and chunk,
// and so the summation order of the norm, do not depend on the order).
__device__ __forceinline__ int fused_xcd_bid(int w, int nb, const FusedGrid& G) {
  if (G.order == 0 || (nb & 7)) return w;
  const int l = (w & 7) * (nb >> 3) + (w >> 3);
  const int nxy = G.gx * G.gy;
  const int bz = l / nxy;
  return (l - bz * nxy) * G.nch + bz;
}

// U, ext: the caller's layout and array of MODE 3 (b out in the solver's
// layout is ext, x_in and b are the caller's) and MODE 4 (ext: the caller's
// solution); unused otherwise.
template <typename T, int TX, int TY, int MODE, int PF>
__global__ void __launch_bounds__(fused_threads(TX, TY))
k_rb_fused(T* __restrict__ xo, const T* __restrict__ xi, const T* __restrict__ b, Lvl<T> L, T omega, FusedGrid G,
           double* __restrict__ partial, Lvl<T> U, T* __restrict__ ext) {
  rb_fused_block<T, TX, TY, MODE, PF>(fused_xcd_bid(blockIdx.x, gridDim.x, G), xo, xi, b, L, omega, G, partial, U,
                                      ext);
}

// Mode 0 marching each chunk from the top down.  A launch that follows an
// upward one on the same level first reads the planes that launch touched
// last, which the memory-side cache still holds (a one-wave grid of blocks
// marches all chunks at once): FP64 256^3 streams of the smoother's shape
// 66 against 81 us per launch, FP32 512^3 300 against 312 us.
template <typename T, int TX, int TY, int MODE, int PF>
__global__ void __launch_bounds__(fused_threads(TX, TY))
k_rb_fused_down(T* __restrict__ xo, const T* __restrict__ xi, const T* __restrict__ b, Lvl<T> L, T omega,
                FusedGrid G, double* __restrict__ partial, Lvl<T> U, T* __restrict__ ext) {
  rb_fused_block<T, TX, TY, MODE, PF, true>(fused_xcd_bid(blockIdx.x, gridDim.x, G), xo, xi, b, L, omega, G, partial,
                                            U, ext);
}

}  // namespace pepsmg
