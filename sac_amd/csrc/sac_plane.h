// sac_plane.h -- the x and y work of an RK stage on 3-D grids in one pass
// over 2-D tiles of a z-plane staged in LDS.
//
// The work is that of the x and y line tiles (sac_line.h) with the same
// arithmetic per value: for each mesh cell the 4th-order flux divergence
// along x and y with the background energy terms (centdiff4,
// src/vaccd.t:35-70), setnu for the 2*ndim+1 variables of each axis
// (src/vacusr.viscosity.t:224-455), the hyperdiffusion line terms and the
// cross-term factors (src/vacusr.viscosity.t:2-221).  Because a tile holds
// both in-plane directions, the cross terms between x and y are
// differentiated here too:
//   m_1 += d_x(rho_t Tm_{1,0}), e += d_x(m_1 Tm_{1,0}), b_1 -= d_x(Tb_{1,0}), e -= d_x(b_1 Tb_{1,0})
//   m_0 += d_y(rho_t Tm_{0,1}), e += d_y(m_0 Tm_{0,1}), b_0 -= d_y(Tb_{0,1}), e -= d_y(b_0 Tb_{0,1})
// (central differences, zero at a non-periodic edge, T = 0 outside the
// mesh), so T_{0,1} and T_{1,0} never leave the CU.  The factors T_{0,2} and
// T_{1,2} go to memory for the z line kernel, which differentiates them
// along z; the rate R is written fresh (the z kernel adds to it).
//
// Schedule.  A block of 256 threads owns a 16 x 16 tile of one z-plane and
// loads the 22 x 22 points around it (3 halo points per side) once: 13
// fields per point, coalesced along x.  (Marching a block through 4-16
// z-planes with the next plane's points loaded during the current one's
// work was 40-55 % slower: the prefetched points and the addresses
// hoisted out of the plane loop need more registers than a CU's two blocks
// can hold.  Evaluating each face ratio once per face instead of once per
// cell that uses it removed 15 % of the instructions but added two
// barriers per axis and was no faster.)  Every point is evaluated once --
// setnu's q values (rho', thermal energy, v, b), rho_t and m -- into LDS, and
// the flux along x (along y) where the x (y) work reads it.  Then every
// thread assembles its cell: the x work from the points of its row, the y
// work from the points of its column (the y fluxes replace the x fluxes in
// one LDS buffer between the two).  The factors T_{0,1} (T_{1,0}) of the
// tile's two neighbouring rows (columns) come from 32 more cells' setnu of
// v_1 and b_1 (v_0 and b_0).
//
// Against separate x and y line kernels this reads the 13 input fields once
// instead of twice and drops the rate accumulator's round trip between
// them (8 fields read and written) and the memory round trip of T_{0,1},
// T_{1,0}.
//
// Grid edges follow the line tiles: a point beyond the grid along x (y) is
// the mirror cell 6 inward (src/vacusr.viscosity.t:378-400), read only by
// the x (y) work of cells in the grid.
#pragma once
#include "sac_device.h"

namespace sac {
namespace plane {

constexpr int TX = 16, TY = 16, HALO = 3;
constexpr int PX = TX + 2 * HALO, PY = TY + 2 * HALO;  // 22 x 22 points
constexpr int NT = TX * TY;
constexpr int NPTS = PX * PY;
// per point values: q of setnu (rho', thermal energy, v0, v1, v2, b0, b1,
// b2), rho_t, m0, m1, m2
constexpr int Q0 = 0, RT = 8, M0 = 9, NV = 12;

// LDS layout.  The flux buffer holds the x fluxes at columns -2..17 of rows
// 0..15, then the y fluxes at columns 0..15 of rows -2..17.  In FP64, 16
// lanes read one tile row of 16 doubles (128 bytes: every bank once) and
// the rows keep their natural widths.  In FP32 a half-wave (32 lanes, served
// together) reads two tile rows; they are rows r and r+8 (see cell_row),
// and every row pitch is 2 (mod 4) floats, so the two rows' reads fall
// 8 * pitch = 16 (mod 32) banks apart and never share a bank.  (Rows r and
// r+1 at the pitch of 22 shared six banks: two cycles per read.)  The
// FP32 layout stays under 40 KB, four blocks per CU.
template <typename C>
struct Lay {
  static constexpr bool F32 = sizeof(C) == 4;
  static constexpr int FX = F32 ? TX + 6 : TX + 4;  // x-flux rows
  static constexpr int FY = F32 ? TX + 2 : TX;      // y-flux rows
  static constexpr int FN = (TY * FX > (TY + 4) * FY) ? TY * FX : (TY + 4) * FY;
  static constexpr int T01 = F32 ? TX + 2 : TX;     // rows of tx01
};

template <typename C>
struct Smem {
  using L = Lay<C>;
  C pv[NV][PY][PX];
  C F[8][L::FN];
  C tx01[2][TY + 2][L::T01];  // Tm_{0,1}, Tb_{0,1} of rows -1..16
  C ty10[2][TY][TX + 2];      // Tm_{1,0}, Tb_{1,0} of columns -1..16
};

// tile row of thread tid: tid / 16 in FP64; in FP32 lanes 0-15, 16-31,
// 32-47, 48-63 of wave w take rows 2w, 2w+8, 2w+1, 2w+9
template <typename C>
__device__ __forceinline__ int cell_row(int tid) {
  if constexpr (Lay<C>::F32)
    return 2 * (tid >> 6) + ((tid >> 5) & 1) + 8 * ((tid >> 4) & 1);
  else
    return tid / TX;
}

__host__ __device__ inline long long nblocks(const GridDesc& g) {
  const long long ntx = (g.Mhi[0] + 1 + TX - 1) / TX;  // tiles from grid column 0
  const long long nty = (g.Mhi[1] - g.Mlo[1] + 1 + TY - 1) / TY;
  return ntx * nty * (g.Mhi[2] - g.Mlo[2] + 1);
}

// One cell's line terms along axis K (0 or 1) from the staged plane, as in
// line_tile: flux divergence, background energy terms, setnu face ratios,
// hyperdiffusion line terms; returns the factors T_{K,l} (l != K, l < 3) in
// Tm[]/Tb[] (index l).  fget(n, d) is the flux of variable n at offset d
// along K, qget(p, d) the q value, rget(d) rho_t, mget(l, d) m_l.
template <int K, typename C, typename FG, typename QG, typename RG, typename MG>
__device__ __forceinline__ void cell_axis(C (&R)[8], const Pt<3, C>& pbg, double gamma, int c,
                                          int MloK, int MhiK, bool zlo, bool zhi, double dxd,
                                          C ctk, int visc_on, bool valid, C& numax, FG fget,
                                          QG qget, RG rget, MG mget, C (&Tm)[3], C (&Tb)[3]) {
  constexpr int NDIR = 3, NDIM = 3, NE = 8;
  const C dxk = (C)dxd, idx = (C)(1.0 / dxd);
  {
    const bool fluxc = c >= MloK + 2 && c <= MhiK - 2;
    const C inv12 = idx / (C)12;
#pragma unroll
    for (int n = 0; n < NE; n++) {
      C E1, e2;
      const C D1 = two_diff(fget(n, 1), fget(n, -1), E1);
      const C D2 = two_diff(fget(n, 2), fget(n, -2), e2);
      R[n] = ((D2 - (C)8 * D1) + (e2 - (C)8 * E1)) * inv12;
    }
    const C pTb0 = (C)ptotal_bg<NDIR, C>(pbg, gamma);
    const C bgk = (C)pbg.bg[K];
    C re = 0;
#pragma unroll
    for (int l = 0; l < NDIM; l++) {
      const C dv1 = qget(2 + l, 1) - qget(2 + l, -1);
      const C dv2 = qget(2 + l, 2) - qget(2 + l, -2);
      const C Dv = -(dv2 - (C)8 * dv1) * inv12;
      if (l == K) re -= Dv * pTb0;
      re += bgk * (C)pbg.bg[l] * Dv;
    }
    R[1 + NDIR] += re;
#pragma unroll
    for (int n = 0; n < NE; n++) R[n] = fluxc ? R[n] : (C)0;
  }
  if (visc_on) {
#pragma clang fp contract(off)
    const bool z1 = (c == MloK && zlo) || (c == MhiK && zhi);
    const C rtm = rget(-1), rt0 = rget(0), rtp = rget(1);
    const C rhoL = (rt0 + rtm) * (C)0.5, rhoR = (rtp + rt0) * (C)0.5;
    // setnu variables of axis K in the order of the line tiles: rho',
    // thermal energy, v_0..v_2, then b_l for l != K
#pragma unroll
    for (int p = 0; p < 7; p++) {
      int qi = p;
      int lb = -1;
      if (p >= 5) {
        const int n = p - 5;
        lb = (n < K) ? n : n + 1;
        qi = 5 + lb;
      }
      C q[7];
#pragma unroll
      for (int s = 0; s < 7; s++) q[s] = qget(qi, s - 3);
      C ratL, ratR;
      face_pair<C>(q, ratL, ratR);
      const C ch = (p >= 2 && p < 5) ? (C)0.4 : (C)0.04;
      const C Kc = ctk * ch;
      const C nuL = Kc * ratL * dxk;
      const C nuR = Kc * ratR * dxk;
      numax = valid ? fmax(numax, fmax(nuR, nuL)) : numax;
      const C qm = q[2], q0 = q[3], qp = q[4];
      const C gL = z1 ? (C)0 : (q0 - qm) * idx;
      const C gR = z1 ? (C)0 : (qp - q0) * idx;
      const C tL = nuL * gL, tR = nuR * gR;
      if (p == 0) {
        R[0] += (tR - tL) * idx;
      } else if (p == 1) {
        R[1 + NDIR] += (tR - tL) * idx;
      } else if (p < 5) {
        const int l = p - 2;
        const C m0 = mget(l, 0);
        const C mL = (m0 + mget(l, -1)) * (C)0.5;
        const C mR = (mget(l, 1) + m0) * (C)0.5;
        const C mult = (l == K) ? (C)1 : (C)0.5;  // the i==k branch runs twice for l==k
        R[1 + l] += mult * (rhoR * tR - rhoL * tL) * idx;
        R[1 + NDIR] += mult * (mR * tR - mL * tL) * idx;
        if (l != K) {
          const C gC = z1 ? (C)0 : (qp - qm) * idx * (C)0.5;
          Tm[l] = gC * (nuL + nuR) * (C)0.25;
        }
      } else {
        const C BL = (q0 + qm) * (C)0.5, BR = (qp + q0) * (C)0.5;
        R[2 + NDIR + lb] += (tR - tL) * idx;
        R[1 + NDIR] += (tR * BR - tL * BL) * idx;
        const C gC = z1 ? (C)0 : (qp - qm) * idx * (C)0.5;
        Tb[lb] = gC * (nuL + nuR) * (C)0.5;
      }
    }
  }
}

// The factor T_{K,l} (Tm, Tb) of one cell from the setnu of v_l and b_l
// along K alone (the halo rows and columns of a tile); the same arithmetic
// as cell_axis.
template <typename C, typename QG>
__device__ __forceinline__ void cell_cross_factor(int l, int c, int MloK, int MhiK, bool zlo,
                                                  bool zhi, double dxd, C ctk, QG qget, C& Tm,
                                                  C& Tb) {
#pragma clang fp contract(off)
  const C dxk = (C)dxd, idx = (C)(1.0 / dxd);
  const bool z1 = (c == MloK && zlo) || (c == MhiK && zhi);
#pragma unroll
  for (int which = 0; which < 2; which++) {
    const int qi = (which == 0) ? 2 + l : 5 + l;
    C q[7];
#pragma unroll
    for (int s = 0; s < 7; s++) q[s] = qget(qi, s - 3);
    const C Kc = ctk * ((which == 0) ? (C)0.4 : (C)0.04);
    C ratL, ratR;
    face_pair<C>(q, ratL, ratR);
    const C nuL = Kc * ratL * dxk;
    const C nuR = Kc * ratR * dxk;
    const C gC = z1 ? (C)0 : (q[4] - q[2]) * idx * (C)0.5;
    if (which == 0)
      Tm = gC * (nuL + nuR) * (C)0.25;
    else
      Tb = gC * (nuL + nuR) * (C)0.5;
  }
}

template <typename S, typename C>
__device__ __forceinline__ void plane_tile(const AxisArgs& a, Smem<C>& sm, C& numax) {
  constexpr int NDIR = 3, NDIM = 3, NE = 8;
  const GridDesc& g = a.g;
  const long long fs = g.fs;
  const S* __restrict__ Wf = (const S*)a.wct;
  const double* __restrict__ Bg = a.bg;
  const double gamma = a.gamma;
  const int tid = threadIdx.x;
  using LY = Lay<C>;
  const int tx = tid % TX, ty = cell_row<C>(tid);

  // block -> tile: each XCD takes one contiguous range of tiles, so the
  // halos a tile shares with its neighbours are fetched into one L2
  long long b = xcd_block(a.xcd);
  if (b >= nblocks(g)) return;
  const int ntx = (g.Mhi[0] + 1 + TX - 1) / TX;
  const int nty = (g.Mhi[1] - g.Mlo[1] + 1 + TY - 1) / TY;
  const int x0 = (int)(b % ntx) * TX;
  b /= ntx;
  const int y0 = g.Mlo[1] + (int)(b % nty) * TY;
  const int z = g.Mlo[2] + (int)(b / nty);
  const long long zoff = (long long)z * g.st[2];

  // ---- points: load and evaluate ----
  C Fy[2][NE];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int i = tid + k * NT;
    if (i >= NPTS) break;
    const int ix = i % PX, iy = i / PX;
    int gx = x0 - HALO + ix, gy = y0 - HALO + iy;
    // mirror layer beyond the grid (used only along its own axis)
    if (gx < 0) gx = g.Mlo[0] + 3;
    if (gx > g.G[0] - 1) gx = g.Mhi[0] - 3;
    if (gy < 0) gy = g.Mlo[1] + 3;
    if (gy > g.G[1] - 1) gy = g.Mhi[1] - 3;
    Pt<NDIR, C> pt;
    load_pt<NDIR, S, C>(Wf, Bg, fs, zoff + (long long)gy * g.st[1] + gx, pt);
    const bool in_fx = ix >= 1 && ix <= PX - 2 && iy >= HALO && iy < HALO + TY;
    const bool in_fy = iy >= 1 && iy <= PY - 2 && ix >= HALO && ix < HALO + TX;
    if (in_fx) {
      C F[NE], v[NDIM];
      flux_pt<NDIR, NDIM, C>(pt, 0, gamma, F, v);
#pragma unroll
      for (int n = 0; n < NE; n++) sm.F[n][(iy - HALO) * LY::FX + ix - 1] = F[n];
    }
    if (in_fy) {
      C v[NDIM];
      flux_pt<NDIR, NDIM, C>(pt, 1, gamma, Fy[k], v);
    }
    const double rtd = (double)pt.rho + pt.rhob;
    sm.pv[RT][iy][ix] = (C)rtd;
#pragma unroll
    for (int l = 0; l < NDIM; l++) sm.pv[M0 + l][iy][ix] = pt.m[l];
    {
#pragma clang fp contract(off)
      const double inv = recip<C>(rtd);
      double sb = 0.0, smm = 0.0;
#pragma unroll
      for (int cd = 0; cd < NDIR; cd++) {
        const double mc = (double)pt.m[cd], bc = (double)pt.b[cd];
        sb += bc * bc;
        smm += mc * mc;
      }
      sm.pv[Q0 + 0][iy][ix] = pt.rho;
      sm.pv[Q0 + 1][iy][ix] = (C)((double)pt.e - 0.5 * (sb + smm * inv));
#pragma unroll
      for (int l = 0; l < NDIM; l++) sm.pv[Q0 + 2 + l][iy][ix] = pt.m[l] * (C)inv;
#pragma unroll
      for (int l = 0; l < NDIR; l++) sm.pv[Q0 + 5 + l][iy][ix] = pt.b[l];
    }
  }

  // ---- the cell ----
  const int cx = x0 + tx, cy = y0 + ty;
  const bool valid = cx >= g.Mlo[0] && cx <= g.Mhi[0] && cy <= g.Mhi[1];
  const int cxc = (cx > g.G[0] - 1) ? g.G[0] - 1 : cx;
  const int cyc = (cy > g.G[1] - 1) ? g.G[1] - 1 : cy;
  const long long icell = zoff + (long long)cyc * g.st[1] + cxc;
  Pt<NDIR, C> pbg;
  pbg.eb = Bg[icell];
#pragma unroll
  for (int c = 0; c < NDIR; c++) pbg.bg[c] = Bg[(2 + c) * fs + icell];
  const double dxx = a.dx[0][cxc], dxy = a.dx[1][cyc];
  const C ctx = a.visc_on ? (C)red_get(a.ctot_src, 0) : (C)0;
  const C cty = a.visc_on ? (C)red_get(a.ctot_src, 1) : (C)0;
  const bool xlo = !g.per_lo[0], xhi = !g.per_hi[0], ylo = !g.per_lo[1], yhi = !g.per_hi[1];
  const int px = tx + HALO, py = ty + HALO;  // the cell's point
  S* __restrict__ Tout = (S*)a.T;
  __syncthreads();

  // x work
  C R[NE];
  {
    C Tm[3] = {0, 0, 0}, Tb[3] = {0, 0, 0};
    cell_axis<0, C>(
        R, pbg, gamma, cx, g.Mlo[0], g.Mhi[0], xlo, xhi, dxx, ctx, a.visc_on, valid, numax,
        [&](int n, int d) { return sm.F[n][ty * LY::FX + tx + 2 + d]; },
        [&](int p, int d) { return sm.pv[Q0 + p][py][px + d]; },
        [&](int d) { return sm.pv[RT][py][px + d]; },
        [&](int l, int d) { return sm.pv[M0 + l][py][px + d]; }, Tm, Tb);
    if (a.visc_on) {
      sm.tx01[0][ty + 1][tx] = Tm[1];
      sm.tx01[1][ty + 1][tx] = Tb[1];
      if (valid) {
        Tout[(long long)cross_index(NDIM, 0, 2) * fs + icell] = (S)Tm[2];
        Tout[(long long)(cross_index(NDIM, 0, 2) + 1) * fs + icell] = (S)Tb[2];
      }
      // T_{0,1} of the rows below and above the tile
      if (tid < 2 * TX) {
        const int hy = (tid < TX) ? -1 : TY;
        const int hpy = hy + HALO;
        C tm, tb;
        cell_cross_factor<C>(1, cx, g.Mlo[0], g.Mhi[0], xlo, xhi, dxx, ctx,
                             [&](int p, int d) { return sm.pv[Q0 + p][hpy][px + d]; }, tm, tb);
        sm.tx01[0][hy + 1][tx] = tm;
        sm.tx01[1][hy + 1][tx] = tb;
      }
    }
  }
  __syncthreads();  // the x fluxes are read; the buffer takes the y fluxes
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int i = tid + k * NT;
    if (i >= NPTS) break;
    const int ix = i % PX, iy = i / PX;
    if (iy >= 1 && iy <= PY - 2 && ix >= HALO && ix < HALO + TX) {
#pragma unroll
      for (int n = 0; n < NE; n++) sm.F[n][(iy - 1) * LY::FY + ix - HALO] = Fy[k][n];
    }
  }
  __syncthreads();

  // y work
  {
    C Ry[NE];
    C Tm[3] = {0, 0, 0}, Tb[3] = {0, 0, 0};
    cell_axis<1, C>(
        Ry, pbg, gamma, cy, g.Mlo[1], g.Mhi[1], ylo, yhi, dxy, cty, a.visc_on, valid, numax,
        [&](int n, int d) { return sm.F[n][(ty + 2 + d) * LY::FY + tx]; },
        [&](int p, int d) { return sm.pv[Q0 + p][py + d][px]; },
        [&](int d) { return sm.pv[RT][py + d][px]; },
        [&](int l, int d) { return sm.pv[M0 + l][py + d][px]; }, Tm, Tb);
#pragma unroll
    for (int n = 0; n < NE; n++) R[n] = R[n] + Ry[n];
    if (a.visc_on) {
      sm.ty10[0][ty][tx + 1] = Tm[0];
      sm.ty10[1][ty][tx + 1] = Tb[0];
      if (valid) {
        Tout[(long long)cross_index(NDIM, 1, 2) * fs + icell] = (S)Tm[2];
        Tout[(long long)(cross_index(NDIM, 1, 2) + 1) * fs + icell] = (S)Tb[2];
      }
      // T_{1,0} of the columns left and right of the tile
      if (tid < 2 * TY) {
        const int hx = (tid < TY) ? -1 : TX;
        const int hpx = hx + HALO;
        const int hty = tid % TY, hcy = y0 + hty;
        const int hcyc = (hcy > g.G[1] - 1) ? g.G[1] - 1 : hcy;
        C tm, tb;
        cell_cross_factor<C>(0, hcy, g.Mlo[1], g.Mhi[1], ylo, yhi, a.dx[1][hcyc], cty,
                             [&](int p, int d) { return sm.pv[Q0 + p][hty + HALO + d][hpx]; },
                             tm, tb);
        sm.ty10[0][hty][hx + 1] = tm;
        sm.ty10[1][hty][hx + 1] = tb;
      }
    }
  }
  __syncthreads();

  // cross terms between x and y (final_cell's formulas, pairs (l=0, k=1)
  // and (l=1, k=0))
  if (a.visc_on) {
#pragma unroll
    for (int l = 0; l < 2; l++) {
      const int k = 1 - l;
      const int cl = (l == 0) ? cx : cy;
      const bool edge = (l == 0) ? ((cx == g.Mlo[0] && xlo) || (cx == g.Mhi[0] && xhi))
                                 : ((cy == g.Mlo[1] && ylo) || (cy == g.Mhi[1] && yhi));
      const C h = (C)(0.5 / ((l == 0) ? dxx : dxy));
      const bool inP = cl + 1 <= g.Mhi[l], inM = cl - 1 >= g.Mlo[l];
      const int dxp = (l == 0) ? 1 : 0, dyp = (l == 0) ? 0 : 1;
      const C rtp = sm.pv[RT][py + dyp][px + dxp], rtm = sm.pv[RT][py - dyp][px - dxp];
      C Tmp0, Tmm0, Tbp0, Tbm0;
      if (l == 0) {
        Tmp0 = sm.ty10[0][ty][tx + 2];
        Tmm0 = sm.ty10[0][ty][tx];
        Tbp0 = sm.ty10[1][ty][tx + 2];
        Tbm0 = sm.ty10[1][ty][tx];
      } else {
        Tmp0 = sm.tx01[0][ty + 2][tx];
        Tmm0 = sm.tx01[0][ty][tx];
        Tbp0 = sm.tx01[1][ty + 2][tx];
        Tbm0 = sm.tx01[1][ty][tx];
      }
      const C Tmp = inP ? Tmp0 : (C)0, Tmm = inM ? Tmm0 : (C)0;
      const C Tbp = inP ? Tbp0 : (C)0, Tbm = inM ? Tbm0 : (C)0;
      const C mkp = sm.pv[M0 + k][py + dyp][px + dxp], mkm = sm.pv[M0 + k][py - dyp][px - dxp];
      const C bkp = sm.pv[Q0 + 5 + k][py + dyp][px + dxp];
      const C bkm = sm.pv[Q0 + 5 + k][py - dyp][px - dxp];
      const C dm = (rtp * Tmp - rtm * Tmm) * h;
      const C de1 = (mkp * Tmp - mkm * Tmm) * h;
      const C db = (Tbp - Tbm) * h;
      const C de2 = (Tbp * bkp - Tbm * bkm) * h;
      R[1 + k] += edge ? (C)0 : dm;
      R[1 + NDIR] += edge ? (C)0 : de1;
      R[2 + NDIR + k] -= edge ? (C)0 : db;
      R[1 + NDIR] -= edge ? (C)0 : de2;
    }
  }
  if (valid) {
    S* __restrict__ Racc = (S*)a.R;
#pragma unroll
    for (int n = 0; n < NE; n++) Racc[n * fs + icell] = (S)R[n];
  }
}

}  // namespace plane
}  // namespace sac
