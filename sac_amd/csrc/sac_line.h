// sac_line.h -- the per-axis work of an RK stage on 2-D and 3-D grids as
// LDS-staged line tiles.
//
// The work is that of axis_group (sac_device.h): for the cells of lines
// along axis K, the 4th-order flux divergence along K with the background
// energy terms (centdiff4, src/vaccd.t:35-70), setnu for the 2*ndim+1
// variables of axis K (src/vacusr.viscosity.t:224-455), the hyperdiffusion
// line terms and the cross-term factors T (src/vacusr.viscosity.t:2-221).
// The arithmetic of each value is the same; the schedule differs.
//
// Schedule.  A block of 256 threads owns W neighbouring lines and marches
// along K in chunks of CH points (16 x 16 along y and z; along x 4 rows of
// 64 points in FP64, 2 rows of 128 in FP32 storage; see Tile).  Every
// thread loads ONE new point per chunk (13 fields, coalesced along x),
// evaluates everything that depends on that point alone (the flux along K,
// rho_t, momenta and the seven setnu q values, which include the
// velocities) once, and stores it in a ring of CH+6 rows in LDS.  After a
// barrier the thread
// assembles the cell three points behind its point from the ring: the
// flux differences of points c-2..c+2 and the face ratios of the 7-point q
// window c-3..c+3.  Each point is loaded from HBM and evaluated once per
// axis, instead of once per neighbour that uses it.
//
// Ring layout: [value][row][line] for K > 0 (a wave reads 4 rows x 16
// consecutive lines, free of LDS bank conflicts), [value][line][row] for
// K = 0 (a wave reads 64 consecutive ring rows of one line).
//
// Grid edges along K follow axis_group: the q layer beyond the grid is the
// mirror of the cell 6 inward (src/vacusr.viscosity.t:378-400); points past
// it are loaded from that same in-grid cell and never reach a mesh cell's
// result.
#pragma once
#include "sac_device.h"

namespace sac {
namespace line {

constexpr int HALO = 3;  // reach of the setnu stencil
constexpr int NT = 256;  // threads per block

// Tile shape per axis and storage type: CH points per chunk along K, W
// lines per block (CH * W = NT).  Along x (K = 0) a chunk is a run of the
// row itself; long runs mean fewer barriers and ring refills per cell, and
// at 256^3 the fastest were 64 points x 4 rows in FP64 (0.97 ms against
// 1.28 for 16 x 16) and 128 x 2 in FP32 storage (0.56 against 0.89).  Along
// y and z, 16 lines x 16 points keeps a 128-byte line of each row in one
// tile row (32-point chunks in 512-thread blocks were slower in FP64).
template <int K, typename C>
struct Tile {
  static constexpr int CH = (K == 0) ? (sizeof(C) == 8 ? 64 : 128) : 16;
  static constexpr int W = (K == 0) ? NT / CH : 16;
  static constexpr int T = CH * W;          // threads per block
  static constexpr int RR = CH + 2 * HALO;  // ring rows
};

// Values per point in the ring.  Only values read at a neighbouring point
// are kept: the background terms of a cell (p_Tb, bg) are read from memory by
// the cell's own thread, and v_l = m_l/rho_t is setnu's q of v_l.  At 19
// FP64 values per point (27 before) three x tiles fit in a CU's LDS instead
// of two (the x kernel 14 % faster at 256^3); the y and z tiles stay at two
// blocks, set by their registers.
template <int NDIM, int NDIR, bool DIVB = false>
struct Slots {
  static constexpr int NE = 2 + 2 * NDIR;
  static constexpr int NP = 2 * NDIM + 1;
  static constexpr int F0 = 0;          // flux along K of the NE evolving fields
  static constexpr int RT = F0 + NE;    // rho_t
  static constexpr int M0 = RT + 1;     // m_l, l < NDIM
  static constexpr int Q0 = M0 + NDIM;  // setnu q of the NP variables: rho', thermal
                                        // energy, v_l (l < NDIM), b_l (l != K)
  static constexpr int V0 = Q0 + 2;     // v_l, l < NDIM (the q of v_l)
  static constexpr int BT = Q0 + NP;    // b_K + bg_K (div B of Powell's sources)
  static constexpr int NV = BT + (DIVB ? 1 : 0);
};

// LDS elements of a line tile (the ring)
template <int NDIM, int NDIR, int K, typename C, bool DIVB>
constexpr int lds_elems() {
  return Slots<NDIM, NDIR, DIVB>::NV * Tile<K, C>::RR * Tile<K, C>::W;
}

template <int K, typename C>
__device__ __forceinline__ int at(int v, int row, int col) {
  constexpr int W = Tile<K, C>::W, RR = Tile<K, C>::RR;
  return (K == 0) ? (v * W + col) * RR + row : (v * RR + row) * W + col;
}

// Axes of a block: A carries the W lines of the block, B (3-D only) one
// line set per block index.
template <int NDIM, int K>
struct Axes {
  static constexpr int A = (K == 0) ? 1 : 0;
  static constexpr int B = (NDIM == 3) ? ((K == 2) ? 1 : 2) : -1;
};

// First grid index of the line tiles along A: tiles along x start at the
// row start, so that a tile's 16 values of a row fill whole 128-byte lines
// (rows are padded to a multiple of 16 elements) and no two tiles share a
// line; the ghost columns of the first and last tile stay idle.
template <int A>
__host__ __device__ inline int tile_origin(const GridDesc& g) {
  return (A == 0) ? 0 : g.Mlo[A];
}

// blocks of a launch for segments of `seg` cells along K
template <int NDIM, int K, typename C>
__host__ __device__ inline long long nblocks(const GridDesc& g, int seg) {
  constexpr int A = Axes<NDIM, K>::A, B = Axes<NDIM, K>::B;
  constexpr int W = Tile<K, C>::W;
  const long long nA = (g.Mhi[A] - tile_origin<A>(g) + 1 + W - 1) / W;
  const long long nB = (B >= 0) ? (g.Mhi[B] - g.Mlo[B] + 1) : 1;
  const long long nK = (g.Mhi[K] - g.Mlo[K] + 1 + seg - 1) / seg;
  return nA * nB * nK;
}

// LEARNED: the face ratios of setnu come from a.RF (learned closure, FP32,
// stored at the face's left cell; < 0 where the exact ratio is required).
// blk >= 0 gives the block's logical index directly (patch ensembles, one
// grid per patch); otherwise it follows from blockIdx.x.
template <int NDIM, int NDIR, int K, typename S, typename C, bool LEARNED = false,
          bool DIVB = false>
__device__ __forceinline__ void line_tile(const AxisArgs& a, int seg, C* __restrict__ L,
                                          C& numax, long long blk = -1) {
  using SL = Slots<NDIM, NDIR, DIVB>;
  constexpr int NE = SL::NE, NP = SL::NP;
  constexpr int A = Axes<NDIM, K>::A, B = Axes<NDIM, K>::B;
  constexpr int CH = Tile<K, C>::CH, W = Tile<K, C>::W, RR = Tile<K, C>::RR;
  const GridDesc& g = a.g;
  const long long fs = g.fs, sk = g.st[K];
  const S* __restrict__ Wf = (const S*)a.wct;
  const double* __restrict__ Bg = a.bg;
  const double gamma = a.gamma;
  const int GK = g.G[K], MloK = g.Mlo[K], MhiK = g.Mhi[K];

  const int tid = threadIdx.x;
  const int ik = (K == 0) ? (tid % CH) : (tid / W);  // position in the chunk
  const int iw = (K == 0) ? (tid / CH) : (tid % W);  // line of the block

  // block -> (line tile along A, line along B, segment along K).  Workgroups
  // are dealt to the 8 XCDs in turn (block b to XCD b % 8); for tiles along x
  // the logical index pairs tiles 2m and 2m+1, whose rows share 128-byte
  // lines when a tile row is shorter than a line (FP32 storage), on one XCD
  // and in consecutive rounds, so the line is fetched into one L2.  The
  // launch is padded to a multiple of 16 blocks; the padding blocks exit.
  long long b = blockIdx.x;
  if (blk >= 0) {
    b = blk;
  } else if constexpr (A == 0) {
    const long long x = b % 8, r = b / 8;
    b = 2 * (x + 8 * (r / 2)) + (r % 2);
  }
  if (b >= nblocks<NDIM, K, C>(g, seg)) return;
  const int nA = (g.Mhi[A] - tile_origin<A>(g) + 1 + W - 1) / W;
  const int bA = (int)(b % nA);
  b /= nA;
  int cB = 0;
  if constexpr (B >= 0) {
    const int nB = g.Mhi[B] - g.Mlo[B] + 1;
    cB = g.Mlo[B] + (int)(b % nB);
    b /= nB;
  }
  const int sIdx = (int)b;
  int cA = tile_origin<A>(g) + bA * W + iw;
  const bool line_ok = cA >= g.Mlo[A] && cA <= g.Mhi[A];
  // an idle line is computed on an in-grid line and never stored
  if (cA > g.Mhi[A]) cA = g.Mhi[A];
  if (cA < g.Mlo[A]) cA = g.Mlo[A];
  int c3[3] = {0, 0, 0};
  c3[A] = cA;
  if constexpr (B >= 0) c3[B] = cB;
  const long long base = lin(g, c3[0], c3[1], c3[2]);  // line origin (index 0 along K)

  const int cs = MloK + sIdx * seg;                      // first cell of the segment
  const int ce = (cs + seg - 1 < MhiK) ? cs + seg - 1 : MhiK;
  // first point of the ring: three before the first cell.  Along x it is
  // moved down so that the cells of every chunk (three behind its points)
  // start on a multiple of 16 elements: each chunk then writes whole
  // 128-byte lines of R and T, and its loads, three elements earlier, share
  // their partial lines with the neighbouring chunk of the same block
  // (2 % faster than aligning the loads at 256^3).  The extra points only
  // feed cells before the segment, which are idle.
  int ps = cs - HALO;
  if (K == 0) {
    const int q = cs - 2 * HALO;
    ps = ((q >= 0) ? q / 16 * 16 : -((-q + 15) / 16) * 16) + HALO;
  }
  const int nch = (ce + HALO - ps + 1 + CH - 1) / CH;

  const C ctk = a.visc_on ? (C)red_get(a.ctot_src, K) : (C)0;
  S* __restrict__ Racc = (S*)a.R;
  S* __restrict__ Tout = (S*)a.T;

  // global index of the point of chunk j (mirror layer beyond the grid)
  auto point_index = [&](int j) {
    int pK = ps + j * CH + ik;
    if (pK < 0) pK = MloK + 3;       // mirror layer below the grid
    if (pK > GK - 1) pK = MhiK - 3;  // and above it
    return base + (long long)pK * sk;
  };
  // global index of the cell of chunk j (clamped into the grid; only cells
  // of the segment are stored)
  auto cell_clamped = [&](int j) {
    const int c = ps + j * CH + ik - HALO;
    return (c < 0) ? 0 : ((c > GK - 1) ? GK - 1 : c);
  };
  auto cell_index = [&](int j) { return base + (long long)cell_clamped(j) * sk; };
  // Software pipeline: the point of chunk j+1 is loaded while chunk j is
  // evaluated.  The cell's own loads (rate accumulator for K > 0, cell
  // width, and on the last axis the cross-term factors) are issued at the
  // start of chunk j, before the next point's loads, so that waiting for
  // them never waits for those; their latency hides behind the point phase.
  //
  // Last axis: the cross-term factors T_{k,K} (k < K) at the cell's
  // neighbours along K, in the order Tm(c+1), Tm(c-1), Tb(c+1), Tb(c-1).
  constexpr bool kCross = (K == NDIM - 1 && NDIM > 1);
  constexpr int NTC = kCross ? 4 * K : 1;
  const bool do_cross = kCross && a.cross_last && a.visc_on;
  Pt<NDIR, C> nxt;
  load_pt<NDIR, S, C>(Wf, Bg, fs, point_index(0), nxt);

  int rb = 0;  // ring row of the chunk's first point
  for (int j = 0; j < nch; j++) {
    const int o = j * CH + ik;  // point offset from ps
    int r0 = rb + ik;
    if (r0 >= RR) r0 -= RR;
    const Pt<NDIR, C> pt = nxt;
    C Rold[NE];
    const long long icell = cell_index(j);
#pragma unroll
    for (int n = 0; n < NE; n++) Rold[n] = (K > 0) ? (C)Racc[n * fs + icell] : (C)0;
    const double dxd = a.dx[K][cell_clamped(j)];
    // the cell's background: p_Tb (getptotal_bg) and bg_l of the energy terms
    Pt<NDIR, C> pbg;
    pbg.eb = Bg[icell];
#pragma unroll
    for (int c = 0; c < NDIR; c++) pbg.bg[c] = Bg[(2 + c) * fs + icell];
    C Tc[NTC];
    if constexpr (kCross) {
      if (do_cross) {
        const int cc = cell_clamped(j);
        const long long ipc = base + (long long)((cc + 1 <= GK - 1) ? cc + 1 : GK - 1) * sk;
        const long long imc = base + (long long)((cc - 1 >= 0) ? cc - 1 : 0) * sk;
#pragma unroll
        for (int k = 0; k < K; k++) {
          const long long xi = (long long)cross_index(NDIM, k, K) * fs;
          Tc[4 * k + 0] = (C)Tout[xi + ipc];
          Tc[4 * k + 1] = (C)Tout[xi + imc];
          Tc[4 * k + 2] = (C)Tout[xi + fs + ipc];
          Tc[4 * k + 3] = (C)Tout[xi + fs + imc];
        }
      }
    }
    // learned face ratios of the cell's two faces (c-1/2, c+1/2) per variable
    float rfc[LEARNED ? 2 * NP : 1];
    if constexpr (LEARNED) {
      const int cc = cell_clamped(j);
      const long long il = base + (long long)((cc - 1 >= 0) ? cc - 1 : 0) * sk;
#pragma unroll
      for (int p = 0; p < NP; p++) {
        rfc[2 * p] = a.RF[(long long)p * fs + il];
        rfc[2 * p + 1] = a.RF[(long long)p * fs + icell];
      }
    }
    if (j + 1 < nch) load_pt<NDIR, S, C>(Wf, Bg, fs, point_index(j + 1), nxt);
    // ---- the new point: evaluate what depends on it alone ----
    {
      C F[NE], v[NDIM];
      flux_pt<NDIR, NDIM, C>(pt, K, gamma, F, v);
#pragma unroll
      for (int n = 0; n < NE; n++) L[at<K, C>(SL::F0 + n, r0, iw)] = F[n];
#pragma unroll
      for (int l = 0; l < NDIM; l++) L[at<K, C>(SL::M0 + l, r0, iw)] = pt.m[l];
      const double rtd = (double)pt.rho + pt.rhob;
      L[at<K, C>(SL::RT, r0, iw)] = (C)rtd;
      if constexpr (DIVB) L[at<K, C>(SL::BT, r0, iw)] = (C)((double)pt.b[K] + pt.bg[K]);
      {
#pragma clang fp contract(off)
        // q of setnu: rho', thermal energy (in FP64), v_l, b_l (l != K)
        const double inv = recip<C>(rtd);
        double sb = 0.0, sm = 0.0;
#pragma unroll
        for (int cd = 0; cd < NDIR; cd++) {
          const double mc = (double)pt.m[cd], bc = (double)pt.b[cd];
          sb += bc * bc;
          sm += mc * mc;
        }
        L[at<K, C>(SL::Q0 + 0, r0, iw)] = pt.rho;
        L[at<K, C>(SL::Q0 + 1, r0, iw)] = (C)((double)pt.e - 0.5 * (sb + sm * inv));
#pragma unroll
        for (int l = 0; l < NDIM; l++) L[at<K, C>(SL::Q0 + 2 + l, r0, iw)] = pt.m[l] * (C)inv;
#pragma unroll
        for (int n = 0; n < NDIM - 1; n++) {
          const int lb = (n < K) ? n : n + 1;
          L[at<K, C>(SL::Q0 + 2 + NDIM + n, r0, iw)] = pt.b[lb];
        }
      }
    }
    __syncthreads();

    // ---- the cell three points behind: c = ps + o - 3 ----
    {
      const int c = ps + o - HALO;
      const bool valid = line_ok && c >= cs && c <= ce;
      auto row = [&](int d) {  // ring row of point c + d
        int r = r0 - HALO + d;
        return (r < 0) ? r + RR : r;
      };
      const C dxk = (C)dxd, idx = (C)(1.0 / dxd);
      C R[NE];
      {
        const bool fluxc = c >= MloK + 2 && c <= MhiK - 2;
        const int rm2 = row(-2), rm1 = row(-1), rp1 = row(1), rp2 = row(2);
        const C inv12 = idx / (C)12;
#pragma unroll
        for (int n = 0; n < NE; n++) {
          C E1, e2;
          const C D1 = two_diff(L[at<K, C>(SL::F0 + n, rp1, iw)], L[at<K, C>(SL::F0 + n, rm1, iw)], E1);
          const C D2 = two_diff(L[at<K, C>(SL::F0 + n, rp2, iw)], L[at<K, C>(SL::F0 + n, rm2, iw)], e2);
          R[n] = ((D2 - (C)8 * D1) + (e2 - (C)8 * E1)) * inv12;
        }
        const C pTb0 = (C)ptotal_bg<NDIR, C>(pbg, gamma);
        const C bgk = (C)pbg.bg[K];
        C re = 0;
#pragma unroll
        for (int l = 0; l < NDIM; l++) {
          const C dv1 = L[at<K, C>(SL::V0 + l, rp1, iw)] - L[at<K, C>(SL::V0 + l, rm1, iw)];
          const C dv2 = L[at<K, C>(SL::V0 + l, rp2, iw)] - L[at<K, C>(SL::V0 + l, rm2, iw)];
          const C Dv = -(dv2 - (C)8 * dv1) * inv12;
          if (l == K) re -= Dv * pTb0;
          re += bgk * (C)pbg.bg[l] * Dv;
        }
        R[1 + NDIR] += re;
#pragma unroll
        for (int n = 0; n < NE; n++) R[n] = fluxc ? R[n] : (C)0;
        // Powell's div B sources (addsource_divb, src/vacphys.mhd0.t:264-297):
        // div B = sum over the axes of gradient4(b_K + bg_K) (zero on the two
        // outermost cells like the flux), so each axis adds its share
        //   m -= (b+bg) g_K,  b -= m/rho_t g_K,  e -= m.(b+bg)/rho_t g_K
        // with the state of the cell (loaded here: all ndir components)
        if (DIVB && valid && fluxc) {
          const C D1b = L[at<K, C>(SL::BT, rp1, iw)] - L[at<K, C>(SL::BT, rm1, iw)];
          const C D2b = L[at<K, C>(SL::BT, rp2, iw)] - L[at<K, C>(SL::BT, rm2, iw)];
          const C gK = -((D2b - (C)8 * D1b) * inv12);
          const double rt = (double)Wf[icell] + Bg[fs + icell];
          double sm = 0.0;
#pragma unroll
          for (int cd = 0; cd < NDIR; cd++) {
            const double m = (double)Wf[(1 + cd) * fs + icell];
            const double bt = (double)Wf[(2 + NDIR + cd) * fs + icell] + Bg[(2 + cd) * fs + icell];
            R[1 + cd] -= (C)bt * gK;
            R[2 + NDIR + cd] -= (C)(m / rt) * gK;
            sm = (cd == 0) ? m * bt : sm + m * bt;
          }
          R[1 + NDIR] -= (C)(sm / rt) * gK;
        }
      }
      // The last axis runs after the others have written their cross-term
      // factors T_{k,K}: their derivatives along K (central, zero at a
      // non-periodic edge of K, T = 0 outside the mesh) are added here from
      // the ring instead of in the combining kernel:
      //   m_k += d_K(rho_t Tm_{k,K}), e += d_K(m_k Tm_{k,K}),
      //   b_k -= d_K(Tb_{k,K}),       e -= d_K(b_k Tb_{k,K})
      if constexpr (kCross) {
        const bool z1 = (c == MloK && !g.per_lo[K]) || (c == MhiK && !g.per_hi[K]);
        if (do_cross && valid && !z1) {
          const int rp1 = row(1), rm1 = row(-1);
          const C h = (C)(0.5 / dxd);
          const bool inP = c + 1 <= MhiK, inM = c - 1 >= MloK;
          const C rtp = L[at<K, C>(SL::RT, rp1, iw)], rtm = L[at<K, C>(SL::RT, rm1, iw)];
#pragma unroll
          for (int k = 0; k < K; k++) {
            const C Tmp = inP ? Tc[4 * k + 0] : (C)0, Tmm = inM ? Tc[4 * k + 1] : (C)0;
            const C Tbp = inP ? Tc[4 * k + 2] : (C)0, Tbm = inM ? Tc[4 * k + 3] : (C)0;
            const C mkp = L[at<K, C>(SL::M0 + k, rp1, iw)], mkm = L[at<K, C>(SL::M0 + k, rm1, iw)];
            const C bkp = L[at<K, C>(SL::Q0 + 2 + NDIM + k, rp1, iw)];
            const C bkm = L[at<K, C>(SL::Q0 + 2 + NDIM + k, rm1, iw)];
            R[1 + k] += (rtp * Tmp - rtm * Tmm) * h;
            R[1 + NDIR] += (mkp * Tmp - mkm * Tmm) * h;
            R[2 + NDIR + k] -= (Tbp - Tbm) * h;
            R[1 + NDIR] -= (Tbp * bkp - Tbm * bkm) * h;
          }
        }
      }
      if (a.visc_on) {
#pragma clang fp contract(off)
        const bool z1 = (c == MloK && !g.per_lo[K]) || (c == MhiK && !g.per_hi[K]);
        const int rw[7] = {row(-3), row(-2), row(-1), row(0), row(1), row(2), row(3)};
        const C rtm = L[at<K, C>(SL::RT, rw[2], iw)], rt0 = L[at<K, C>(SL::RT, rw[3], iw)],
                rtp = L[at<K, C>(SL::RT, rw[4], iw)];
        const C rhoL = (rt0 + rtm) * (C)0.5, rhoR = (rtp + rt0) * (C)0.5;
#pragma unroll
        for (int p = 0; p < NP; p++) {
          C q[7];
#pragma unroll
          for (int s = 0; s < 7; s++) q[s] = L[at<K, C>(SL::Q0 + p, rw[s], iw)];
          const C ch = (p >= 2 && p < 2 + NDIM) ? (C)0.4 : (C)0.04;
          const C Kc = ctk * ch;
          C ratL, ratR;
          if constexpr (LEARNED) {
            ratL = (rfc[2 * p] >= 0.f) ? (C)rfc[2 * p] : face_ratio<C>(q);
            ratR = (rfc[2 * p + 1] >= 0.f) ? (C)rfc[2 * p + 1] : face_ratio<C>(q + 1);
          } else {
            face_pair<C>(q, ratL, ratR);
          }
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
          } else if (p < 2 + NDIM) {
            const int l = p - 2;
            const C m0 = L[at<K, C>(SL::M0 + l, rw[3], iw)];
            const C mL = (m0 + L[at<K, C>(SL::M0 + l, rw[2], iw)]) * (C)0.5;
            const C mR = (L[at<K, C>(SL::M0 + l, rw[4], iw)] + m0) * (C)0.5;
            const C mult = (l == K) ? (C)1 : (C)0.5;  // the i==k branch runs twice for l==k
            R[1 + l] += mult * (rhoR * tR - rhoL * tL) * idx;
            R[1 + NDIR] += mult * (mR * tR - mL * tL) * idx;
            if (l != K && valid) {
              const C gC = z1 ? (C)0 : (qp - qm) * idx * (C)0.5;
              Tout[(long long)cross_index(NDIM, K, l) * fs + icell] = (S)(gC * (nuL + nuR) * (C)0.25);
            }
          } else {
            const int n = p - 2 - NDIM;
            const int lb = (n < K) ? n : n + 1;
            const C BL = (q0 + qm) * (C)0.5, BR = (qp + q0) * (C)0.5;
            R[2 + NDIR + lb] += (tR - tL) * idx;
            R[1 + NDIR] += (tR * BR - tL * BL) * idx;
            if (valid) {
              const C gC = z1 ? (C)0 : (qp - qm) * idx * (C)0.5;
              Tout[(long long)(cross_index(NDIM, K, lb) + 1) * fs + icell] =
                  (S)(gC * (nuL + nuR) * (C)0.5);
            }
          }
        }
      }
      if (valid) {
        if (K == 0) {
#pragma unroll
          for (int n = 0; n < NE; n++) Racc[n * fs + icell] = (S)R[n];
        } else {
#pragma unroll
          for (int n = 0; n < NE; n++) Racc[n * fs + icell] = (S)(Rold[n] + R[n]);
        }
      }
    }
    __syncthreads();  // the next chunk overwrites ring rows read above
    rb += CH;
    if (rb >= RR) rb -= RR;
  }
}

}  // namespace line
}  // namespace sac
