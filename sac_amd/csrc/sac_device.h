// sac_device.h -- device-side building blocks of the SAC perturbation-MHD
// stage shared by the grid kernels (sac_kernels.hip) and the batched tube
// kernels (sac_batch.hip): point state, fluxes, cmax, setnu face ratios,
// the per-axis line-group work, the per-cell combining update and the
// ghost-value chain.
//
// Every kernel is a template over the storage type S of the evolving fields
// (and of the rate and cross-term scratch) and the compute type C of the
// flux and hyperdiffusion arithmetic:
//   FP64:  S = C = double;
//   mixed: S = C = float, with FP64 kept for the background fields, the
//          pressure recovery from the split total energy (getptotal and the
//          thermal energy of setnu), the register update, cmax and every
//          reduction; the flux differences of gradient4 are compensated.
//
// Reference routines restated here:
//   getflux            src/vacphys.mhd0.t:147-222
//   getptotal(_bg)     src/vacphys0.t.mhd:186-224
//   getcmax            src/vacphys0.t.mhd:77-112
//   gradient4          src/vacgrid.t:665-712
//   centdiff4          src/vaccd.t:7-78
//   addsource_grav     src/vacusr.gravity.t:24-81
//   addsource_visc     src/vacusr.viscosity.t:2-221
//   setnu              src/vacusr.viscosity.t:224-455
//   gradient1/1L/1R    src/vacusr.viscosity.t:547-675
//   getboundary        src/vacgrid.t:195-512
#pragma once
#include "sac_kernels.h"


// keeps the machine scheduler from hoisting the next phase's loads above
// the current phase's arithmetic, which bounds the number of live values
#define SCHED_FENCE()                  \
  do {                                 \
    asm volatile("" ::: "memory");     \
    __builtin_amdgcn_sched_barrier(0); \
  } while (0)

namespace sac {

// ---------------------------------------------------------------------
// precision helpers

// a/b for 0 <= a <= 6b, b > 0 (the setnu ratio md3/md1): IEEE division in
// FP64; in FP32 a reciprocal-multiply, with both operands scaled by 2^100
// (exact) when b is below 2^-100 so that the reciprocal of a denormal never
// overflows
__device__ __forceinline__ double cdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float cdiv(float a, float b) {
  const float s = (b < 0x1p-100f) ? 0x1p100f : 1.0f;
  return (a * s) * __builtin_amdgcn_rcpf(b * s);
}

// 1/x in FP64 for the pressure recovery: IEEE division in FP64 mode; in
// mixed mode the FP32 reciprocal refined by one Newton step in FP64
// (relative error ~1e-14)
template <typename C>
__device__ __forceinline__ double recip(double x);
template <>
__device__ __forceinline__ double recip<double>(double x) {
  return 1.0 / x;
}
template <>
__device__ __forceinline__ double recip<float>(double x) {
  const double r = (double)__builtin_amdgcn_rcpf((float)x);
  return fma(fma(-x, r, 1.0), r, r);
}

// a-b with its rounding error (FP64 mode: error 0, folded away)
__device__ __forceinline__ double two_diff(double a, double b, double& err) {
  err = 0.0;
  return a - b;
}
__device__ __forceinline__ float two_diff(float a, float b, float& err) {
  const float s = a - b;
  const float bb = s - a;
  err = (a - (s - bb)) - (b + bb);
  return s;
}

// ---------------------------------------------------------------------
// point state: perturbation (evolving) in C, static background in FP64

template <int NDIR, typename C>
struct Pt {
  C rho, m[NDIR], e, b[NDIR];
  double eb, rhob, bg[NDIR];
};

template <int NDIR, typename S, typename C>
__device__ __forceinline__ void load_pt(const S* __restrict__ W, const double* __restrict__ B,
                                        long long fs, long long i, Pt<NDIR, C>& p) {
  p.rho = (C)W[i];
#pragma unroll
  for (int c = 0; c < NDIR; c++) p.m[c] = (C)W[(1 + c) * fs + i];
  p.e = (C)W[(1 + NDIR) * fs + i];
#pragma unroll
  for (int c = 0; c < NDIR; c++) p.b[c] = (C)W[(2 + NDIR + c) * fs + i];
  p.eb = B[i];
  p.rhob = B[fs + i];
#pragma unroll
  for (int c = 0; c < NDIR; c++) p.bg[c] = B[(2 + c) * fs + i];
}

// perturbation total pressure p_T' in FP64 (getptotal, src/vacphys0.t.mhd:186-205);
// inv = 1/rho_t
template <int NDIR, typename C>
__device__ __forceinline__ double ptotal(const Pt<NDIR, C>& p, double gamma, double inv) {
  double sbbg = 0.0, sb2 = 0.0, sm2 = 0.0;
#pragma unroll
  for (int c = 0; c < NDIR; c++) {
    const double b = (double)p.b[c], m = (double)p.m[c];
    sbbg += b * p.bg[c];
    sb2 += b * b;
    sm2 += m * m;
  }
  return (gamma - 1.0) * ((double)p.e - 0.5 * sm2 * inv) - (gamma - 2.0) * (sbbg + 0.5 * sb2);
}

// background total pressure (getptotal_bg, src/vacphys0.t.mhd:208-224)
template <int NDIR, typename C>
__device__ __forceinline__ double ptotal_bg(const Pt<NDIR, C>& p, double gamma) {
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < NDIR; c++) s += p.bg[c] * p.bg[c];
  return (gamma - 1.0) * p.eb - 0.5 * (gamma - 2.0) * s;
}

// full flux (non-transport part of getflux + transport v_d*w) along axis d;
// also returns v_l = m_l/rho_t for l < NV
template <int NDIR, int NV, typename C>
__device__ __forceinline__ void flux_pt(const Pt<NDIR, C>& p, int d, double gamma,
                                        C (&F)[2 + 2 * NDIR], C (&v)[NV]) {
  const double rt = (double)p.rho + p.rhob;
  const double invd = recip<C>(rt);
  const C inv = (C)invd;
  const C vd = p.m[d] * inv;
  F[0] = (C)p.rhob * p.m[d] * inv + vd * p.rho;
  const C pT = (C)ptotal<NDIR, C>(p, gamma, invd);
  C sbgm = 0, sbm = 0;
#pragma unroll
  for (int c = 0; c < NDIR; c++) {
    sbgm += (C)p.bg[c] * p.m[c];
    sbm += p.b[c] * p.m[c];
  }
#pragma unroll
  for (int c = 0; c < NDIR; c++) {
    C f = -(p.b[c] * (C)p.bg[d] + p.b[d] * (C)p.bg[c]) - p.b[c] * p.b[d];
    if (c == d) f = pT + f;
    F[1 + c] = f + vd * p.m[c];
  }
  F[1 + NDIR] = (p.m[d] * pT - p.b[d] * sbgm - (C)p.bg[d] * sbm) * inv +
                (C)p.eb * p.m[d] * inv - p.b[d] * sbm * inv + vd * p.e;
#pragma unroll
  for (int c = 0; c < NDIR; c++) {
    F[2 + NDIR + c] = (c == d) ? (C)0
                               : (-p.m[c] * inv * (p.b[d] + (C)p.bg[d]) +
                                  p.m[d] * inv * (C)p.bg[c] + vd * p.b[c]);
  }
#pragma unroll
  for (int l = 0; l < NV; l++) v[l] = p.m[l] * inv;
}

// fast-mode speed + |v_d| (getcmax, src/vacphys0.t.mhd:77-112) in FP64 for
// every direction d < NDIM; the direction-independent part (c_s^2, c_f^2,
// 1/rho_t) is evaluated once
template <int NDIM, int NDIR>
__device__ __forceinline__ void cmax_all(const Pt<NDIR, double>& p, double gamma,
                                         double (&cm)[NDIM]) {
  const double rt = p.rho + p.rhob;
  const double inv = 1.0 / rt;
  double sm2 = 0.0, sb2 = 0.0, sbbg = 0.0, sbg2 = 0.0, sbt2 = 0.0;
#pragma unroll
  for (int c = 0; c < NDIR; c++) {
    sm2 += p.m[c] * p.m[c];
    sb2 += p.b[c] * p.b[c];
    sbbg += p.b[c] * p.bg[c];
    sbg2 += p.bg[c] * p.bg[c];
    double bt = p.b[c] + p.bg[c];
    sbt2 += bt * bt;
  }
  double pth = 0.5 * sm2 * inv;
  pth = pth + 0.5 * sb2 + sbbg;
  pth = (gamma - 1.0) * (p.e - pth);
  const double cs2 = gamma * (pth + (gamma - 1.0) * (p.eb - 0.5 * sbg2)) * inv;
  const double cf2 = sbt2 * inv + cs2;
  const double cf4 = cf2 * cf2, cs4 = 4.0 * cs2 * inv;
#pragma unroll
  for (int d = 0; d < NDIM; d++) {
    const double bd = p.b[d] + p.bg[d];
    cm[d] = sqrt(0.5 * (cf2 + sqrt(cf4 - cs4 * (bd * bd)))) + fabs(p.m[d] * inv);
  }
}

// ---------------------------------------------------------------------
// block-wide max reduction into uint64 atomics (values are >= 0)

template <int NV>
__device__ __forceinline__ void block_max_atomic(double (&v)[NV], unsigned long long* dst) {
  __shared__ double red[NV][16];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    if (lane == 0) red[k][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int nwave = (blockDim.x + 63) >> 6;
    double x = 0.0;
    for (int j = 0; j < nwave; j++) x = fmax(x, red[threadIdx.x][j]);
    atomicMax(dst + threadIdx.x, (unsigned long long)__double_as_longlong(x));
  }
}

__device__ __forceinline__ double red_get(const unsigned long long* p, int k) {
  return __longlong_as_double((long long)p[k]);
}

// ---------------------------------------------------------------------
// index helpers

__device__ __forceinline__ void decode_box(long long t, const int (&lo)[3], const int (&ext)[3],
                                           int (&c)[3]) {
  c[0] = lo[0] + (int)(t % ext[0]);
  t /= ext[0];
  c[1] = lo[1] + (int)(t % ext[1]);
  t /= ext[1];
  c[2] = lo[2] + (int)t;
}

__device__ __forceinline__ long long lin(const GridDesc& g, int i0, int i1, int i2) {
  return (long long)i0 * g.st[0] + (long long)i1 * g.st[1] + (long long)i2 * g.st[2];
}

// setnu ratio md3/md1 at the face between q[j] and q[j+1] of a 6-value
// window q[j-2..j+3] (src/vacusr.viscosity.t:413-427)
// The q values arrive rounded, as the reference stores them in tmp_nuI; no
// product may be fused into these differences, or a constant region (all
// q equal) would show rounding residues instead of exact zeros, and the
// ratio there (and maxviscoef, and dt) would depend on them.
template <typename C>
__device__ __forceinline__ C face_ratio(const C* q) {
#pragma clang fp contract(off)
  C d1[5], d3[3];
#pragma unroll
  for (int j = 0; j < 5; j++) d1[j] = fabs(q[j + 1] - q[j]);
#pragma unroll
  for (int j = 0; j < 3; j++) d3[j] = fabs((C)3 * (q[j + 2] - q[j + 1]) - (q[j + 3] - q[j]));
  const C md3 = fmax(fmax(d3[0], d3[1]), d3[2]);
  const C md1 = fmax(fmax(fmax(d1[0], d1[1]), fmax(d1[2], d1[3])), d1[4]);
  return (md1 > (C)0) ? cdiv(md3, md1) : (C)0;
}

// The ratios of both faces of a cell, c-1/2 and c+1/2, from its 7-point
// window q[0..6] (c-3..c+3): equal to face_ratio(q) and face_ratio(q + 1).
// The two windows share four of their five first differences and two of
// their three third differences; the maximum over the shared ones is taken
// once (max is exact, so the grouping does not change the result).
template <typename C>
__device__ __forceinline__ void face_pair(const C* q, C& ratL, C& ratR) {
#pragma clang fp contract(off)
  C s[6], d3[4];
#pragma unroll
  for (int j = 0; j < 6; j++) s[j] = q[j + 1] - q[j];
#pragma unroll
  for (int j = 0; j < 4; j++) d3[j] = fabs((C)3 * s[j + 1] - (q[j + 3] - q[j]));
  const C m1 = fmax(fmax(fabs(s[1]), fabs(s[2])), fmax(fabs(s[3]), fabs(s[4])));
  const C m3 = fmax(d3[1], d3[2]);
  const C md1L = fmax(m1, fabs(s[0])), md1R = fmax(m1, fabs(s[5]));
  const C md3L = fmax(d3[0], m3), md3R = fmax(m3, d3[3]);
  ratL = (md1L > (C)0) ? cdiv(md3L, md1L) : (C)0;
  ratR = (md1R > (C)0) ? cdiv(md3R, md1R) : (C)0;
}

__device__ __forceinline__ int cross_index(int ndim, int k, int l) {
  return 2 * (k * (ndim - 1) + (l < k ? l : l - 1));
}

// Workgroups are dealt to the 8 XCDs of the device in turn; renumbering them
// so that each XCD takes one contiguous range of cells keeps the neighbour
// rows of a line stencil in the same XCD's L2.  The grid is a multiple of 8
// blocks, so the renumbering is a bijection.
__device__ __forceinline__ long long xcd_block(int xcd) {
  const unsigned nb = gridDim.x;
  if (xcd <= 1 || (nb % (unsigned)xcd) != 0) return blockIdx.x;
  const unsigned per = nb / (unsigned)xcd;
  return (long long)(blockIdx.x % (unsigned)xcd) * per + blockIdx.x / (unsigned)xcd;
}

// ---------------------------------------------------------------------
// kernel 1: the work of one axis K for a group of CK consecutive mesh cells
// along K, all of which lies on the K-line through the group:
//   - 4th-order flux divergence along K (centdiff4, src/vaccd.t:35-70)
//     and the background energy terms with derivatives along K,
//   - setnu (src/vacusr.viscosity.t:224-455) for the 2*ndim+1 variables of
//     axis K from the 7-point stencil (mirror layer beyond the grid),
//   - the hyperdiffusion terms along K of addsource_visc (rho, thermal
//     energy, momentum i==k branch, field m==k branch),
//   - the cross-term factors T (diffusion along K) that kernel 2
//     differentiates along the other axes:
//       Tm_{K,l} = d_K v_l * (nuL+nuR)_{m_l}/4,  Tb_{K,l} = d_K b_l * (nuL+nuR)_{b_l}/2.
// The group shares its setnu work: the q values of the CK+6 stencil points
// and the CK+1 face ratios are evaluated once for the CK cells.  The
// contribution is added to the rate accumulator R (written on K=0).

// mesh cells per thread of a line group: 2 along every axis
template <int K>
struct LineGroup {
  static constexpr int CK = 2;
};

// One group of CK consecutive mesh cells along K starting at c[K] (other
// coordinates c[d]) of the grid g; adds the group's maximum nu to numax.
template <int NDIM, int NDIR, int K, int CK, typename S, typename C, bool LEARNED = false>
__device__ __forceinline__ void axis_group(const AxisArgs& a, const GridDesc& g, const int (&c)[3],
                                           C& numax) {
  constexpr int NE = 2 + 2 * NDIR;
  constexpr int NP = 2 * NDIM + 1;
  const int c0 = c[K];
  const long long fs = g.fs;
  const S* __restrict__ W = (const S*)a.wct;
  const double* __restrict__ B = a.bg;
  const double gamma = a.gamma;
  const long long ic0 = lin(g, c[0], c[1], c[2]);
  const long long sk = g.st[K];
  const int GK = g.G[K], MloK = g.Mlo[K], MhiK = g.Mhi[K];
  // per-cell geometry; cells beyond the mesh are computed on clamped
  // indices and never read back
  bool valid[CK], fluxc[CK], z1[CK];
  C dxk[CK], idx[CK];
#pragma unroll
  for (int j = 0; j < CK; j++) {
    const int cj = c0 + j;
    valid[j] = cj <= MhiK;
    fluxc[j] = valid[j] && cj >= MloK + 2 && cj <= MhiK - 2;
    z1[j] = (cj == MloK && !g.per_lo[K]) || (cj == MhiK && !g.per_hi[K]);
    const double dxd = a.dx[K][cj < GK ? cj : GK - 1];
    dxk[j] = (C)dxd;
    idx[j] = (C)(1.0 / dxd);
  }
  C R[CK][NE];

  // ---- flux divergence along K (zero on two layers at each side) ----
  // cell by cell: the four neighbour fluxes of a cell (the neighbours are
  // shared with the next cells through the caches).  Branch-free: a cell
  // outside the flux range is evaluated on an in-grid cell and its result
  // replaced by zero.
#pragma unroll
  for (int j = 0; j < CK; j++) {
    const int cf = (c0 + j > MhiK) ? MhiK : c0 + j;
    const long long ic = ic0 + (long long)(cf - c0) * sk;
    C bgk, bgl[NDIM], pTb0;
    {
      Pt<NDIR, C> p0;
      load_pt<NDIR, S, C>(W, B, fs, ic, p0);
      pTb0 = (C)ptotal_bg<NDIR, C>(p0, gamma);
      bgk = (C)p0.bg[K];
#pragma unroll
      for (int l = 0; l < NDIM; l++) bgl[l] = (C)p0.bg[l];
    }
    const C inv12 = idx[j] / (C)12;
    C Fm[NE], D1[NE], E1[NE], vm[NDIM], dv1[NDIM];
    SCHED_FENCE();
    {
      Pt<NDIR, C> p;
      load_pt<NDIR, S, C>(W, B, fs, ic - sk, p);
      flux_pt<NDIR, NDIM, C>(p, K, gamma, Fm, vm);
    }
    SCHED_FENCE();
    {
      Pt<NDIR, C> p;
      C F[NE], v[NDIM];
      load_pt<NDIR, S, C>(W, B, fs, ic + sk, p);
      flux_pt<NDIR, NDIM, C>(p, K, gamma, F, v);
#pragma unroll
      for (int n = 0; n < NE; n++) D1[n] = two_diff(F[n], Fm[n], E1[n]);
#pragma unroll
      for (int l = 0; l < NDIM; l++) dv1[l] = v[l] - vm[l];
    }
    SCHED_FENCE();
    {
      Pt<NDIR, C> p;
      load_pt<NDIR, S, C>(W, B, fs, ic - 2 * sk, p);
      flux_pt<NDIR, NDIM, C>(p, K, gamma, Fm, vm);
    }
    SCHED_FENCE();
    {
      Pt<NDIR, C> p;
      C F[NE], v[NDIM];
      load_pt<NDIR, S, C>(W, B, fs, ic + 2 * sk, p);
      flux_pt<NDIR, NDIM, C>(p, K, gamma, F, v);
      // -gradient4 = ((F2-F-2) - 8 (F1-F-1)) / (12 dx), with the rounding
      // errors of the two differences added back (zero in FP64)
#pragma unroll
      for (int n = 0; n < NE; n++) {
        C e2;
        const C D2 = two_diff(F[n], Fm[n], e2);
        R[j][n] = ((D2 - (C)8 * D1[n]) + (e2 - (C)8 * E1[n])) * inv12;
      }
      C re = 0;
#pragma unroll
      for (int l = 0; l < NDIM; l++) {
        const C Dv = -((v[l] - vm[l]) - (C)8 * dv1[l]) * inv12;
        if (l == K) re -= Dv * pTb0;
        re += bgk * bgl[l] * Dv;
      }
      R[j][1 + NDIR] += re;
#pragma unroll
      for (int n = 0; n < NE; n++) R[j][n] = fluxc[j] ? R[j][n] : (C)0;
    }
    SCHED_FENCE();
  }

  // ---- hyperdiffusion along K ----
  if (a.visc_on) {
    const C ctk = (C)red_get(a.ctot_src, K);
    constexpr int NQ = CK + 6;  // stencil points at offsets -3..CK+2
    int off[NQ];
    double inv7[NQ];
#pragma unroll
    for (int s = 0; s < NQ; s++) {
      int j = c0 + s - 3;
      if (j < 0) j = MloK + 3;       // mirror layer below the grid
      if (j > GK - 1) j = MhiK - 3;  // and above it
      off[s] = (j - c0) * (int)sk;
      const long long i = ic0 + off[s];
      inv7[s] = recip<C>((double)W[i] + B[fs + i]);
    }
    // rho_t of the cells and their direct neighbours (offsets -1..CK)
    C rtw[CK + 2];
    int offn[CK + 2];
#pragma unroll
    for (int s = 0; s < CK + 2; s++) {
      int j = c0 + s - 1;
      j = j > GK - 1 ? GK - 1 : j;
      offn[s] = (j - c0) * (int)sk;
      const long long i = ic0 + offn[s];
      rtw[s] = (C)((double)W[i] + B[fs + i]);
    }
    // p = 0: rho' ; p = 1: thermal energy ; p = 2+l: v_l ; then b_l (l != K).
    // No FP contraction here: differences of q (face ratios, gradient1L/R)
    // must see the rounded q values, as in the reference.
#pragma unroll
    for (int p = 0; p < NP; p++) {
#pragma clang fp contract(off)
      C q[NQ];
      int lb = -1;
      if (p >= 2 + NDIM) {
        int n = p - 2 - NDIM;
        lb = (n < K) ? n : n + 1;
      }
#pragma unroll
      for (int s = 0; s < NQ; s++) {
        const long long i = ic0 + off[s];
        if (p == 0) {
          q[s] = (C)W[i];
        } else if (p == 1) {
          // thermal energy e - (b^2 + m^2/rho_t)/2, recovered in FP64
          double sb = 0.0, sm = 0.0;
#pragma unroll
          for (int cd = 0; cd < NDIR; cd++) {
            const double mc = (double)W[(1 + cd) * fs + i];
            const double bc = (double)W[(2 + NDIR + cd) * fs + i];
            sb += bc * bc;
            sm += mc * mc;
          }
          q[s] = (C)((double)W[(1 + NDIR) * fs + i] - 0.5 * (sb + sm * inv7[s]));
        } else if (p < 2 + NDIM) {
          q[s] = (C)W[(p - 1) * fs + i] * (C)inv7[s];
        } else {
          q[s] = (C)W[(2 + NDIR + lb) * fs + i];
        }
      }
      const C ch = (p >= 2 && p < 2 + NDIM) ? (C)0.4 : (C)0.04;
      const C Kc = ctk * ch;
      // face f (between offsets f-1 and f) from the window q[f..f+5]
      C ratio[CK + 1];
#pragma unroll
      for (int f = 0; f < CK + 1; f++) {
        // learned closure: the predicted ratio of this face (stored at its
        // left cell), or -1 where the exact ratio is required
        if constexpr (LEARNED) {
          const float rf = a.RF[(long long)p * fs + ic0 + (long long)(f - 1) * sk];
          if (rf >= 0.f)
            ratio[f] = (C)rf;
          else
            ratio[f] = face_ratio<C>(q + f);
        } else {
          ratio[f] = face_ratio<C>(q + f);
        }
      }
#pragma unroll
      for (int j = 0; j < CK; j++) {
        const C nuR = Kc * ratio[j + 1] * dxk[j];
        const C nuL = Kc * ratio[j] * dxk[j];
        numax = valid[j] ? fmax(numax, fmax(nuR, nuL)) : numax;
        const C qm = q[j + 2], q0 = q[j + 3], qp = q[j + 4];
        const C gL = z1[j] ? (C)0 : (q0 - qm) * idx[j];
        const C gR = z1[j] ? (C)0 : (qp - q0) * idx[j];
        const C tL = nuL * gL, tR = nuR * gR;
        const long long icj = ic0 + offn[j + 1];
        if (p == 0) {
          R[j][0] += (tR - tL) * idx[j];
        } else if (p == 1) {
          R[j][1 + NDIR] += (tR - tL) * idx[j];
        } else if (p < 2 + NDIM) {
          const int l = p - 2;
          const C rhoL = (rtw[j + 1] + rtw[j]) * (C)0.5, rhoR = (rtw[j + 2] + rtw[j + 1]) * (C)0.5;
          const S* __restrict__ Wm = W + (1 + l) * fs + ic0;
          const C m0 = (C)Wm[offn[j + 1]];
          const C mL = (m0 + (C)Wm[offn[j]]) * (C)0.5;
          const C mR = ((C)Wm[offn[j + 2]] + m0) * (C)0.5;
          const C mult = (l == K) ? (C)1 : (C)0.5;  // the i==k branch runs twice for l==k
          R[j][1 + l] += mult * (rhoR * tR - rhoL * tL) * idx[j];
          R[j][1 + NDIR] += mult * (mR * tR - mL * tL) * idx[j];
          if (l != K) {
            const C gC = z1[j] ? (C)0 : (qp - qm) * idx[j] * (C)0.5;
            ((S*)a.T)[(long long)cross_index(NDIM, K, l) * fs + icj] =
                (S)(gC * (nuL + nuR) * (C)0.25);
          }
        } else {
          const C BL = (q0 + qm) * (C)0.5, BR = (qp + q0) * (C)0.5;
          R[j][2 + NDIR + lb] += (tR - tL) * idx[j];
          R[j][1 + NDIR] += (tR * BR - tL * BL) * idx[j];
          const C gC = z1[j] ? (C)0 : (qp - qm) * idx[j] * (C)0.5;
          ((S*)a.T)[(long long)(cross_index(NDIM, K, lb) + 1) * fs + icj] =
              (S)(gC * (nuL + nuR) * (C)0.5);
        }
      }
      SCHED_FENCE();
    }
  }
  // ---- accumulate ----
  // Stores are unconditional (branches here cost registers): a cell past
  // the mesh end writes R and T of a ghost cell (index clamped to the
  // grid), which nothing reads.
  S* __restrict__ Racc = (S*)a.R;
#pragma unroll
  for (int j = 0; j < CK; j++) {
    const int cj = (c0 + j > GK - 1) ? GK - 1 : c0 + j;
    const long long icj = ic0 + (long long)(cj - c0) * sk;
    if (K == 0) {
#pragma unroll
      for (int n = 0; n < NE; n++) Racc[n * fs + icj] = (S)R[j][n];
    } else {
#pragma unroll
      for (int n = 0; n < NE; n++) Racc[n * fs + icj] = (S)((C)Racc[n * fs + icj] + R[j][n]);
    }
  }
}

// ---------------------------------------------------------------------
// 1-D grids in FP64: the x work of a tile of LT consecutive cells c0..c0+LT-1
// with LDS staging.  The LT+6 points the tile reads (cells -3..+3 around,
// the mirror layer beyond the grid as in axis_group) are loaded and
// evaluated once each: flux and velocity along x (getflux + transport),
// p_Tb and bg_x of the background energy terms, rho_t and m_x of the face
// averages, and setnu's q for rho', the thermal energy and v_x.  Each
// thread then assembles its cell from its neighbours' values with the
// expressions of axis_group<1, NDIR, 0, 1> in the same order, but with each
// stencil point evaluated once instead of up to seven times.  R is stored
// for the mesh cells of the tile; numax gets the thread's largest nu.
namespace line1 {
// values per point: F[NE], v, pTb, bg, rho_t, m, q[3]
template <int NDIR>
constexpr int LV = 2 + 2 * NDIR + 8;
}  // namespace line1

template <int LT, int NDIR>
__device__ __forceinline__ void line1_tile(const GridDesc& g, const double* __restrict__ W,
                                           const double* __restrict__ B,
                                           const double* __restrict__ dx,
                                           double* __restrict__ Rout, double gamma, int visc_on,
                                           double ctk, int c0, double* __restrict__ L,
                                           double& numax) {
  constexpr int NE = 2 + 2 * NDIR, LP = LT + 6;
  const long long fs = g.fs;
  const int GK = g.G[0], MloK = g.Mlo[0], MhiK = g.Mhi[0];
  const int t = (int)threadIdx.x;
  auto at = [&](int v, int i) -> double& { return L[v * LP + i]; };
  for (int i = t; i < LP; i += LT) {
    int pK = c0 - 3 + i;
    if (pK < 0) pK = MloK + 3;       // mirror layer below the grid
    if (pK > GK - 1) pK = MhiK - 3;  // and above it
    Pt<NDIR, double> p;
    load_pt<NDIR, double, double>(W, B, fs, pK, p);
    double F[NE], v[1];
    flux_pt<NDIR, 1, double>(p, 0, gamma, F, v);
#pragma unroll
    for (int n = 0; n < NE; n++) at(n, i) = F[n];
    at(NE, i) = v[0];
    at(NE + 1, i) = ptotal_bg<NDIR, double>(p, gamma);
    at(NE + 2, i) = p.bg[0];
    at(NE + 3, i) = p.rho + p.rhob;
    at(NE + 4, i) = p.m[0];
    {
#pragma clang fp contract(off)
      const double inv = recip<double>(p.rho + p.rhob);
      double sb = 0.0, sm = 0.0;
#pragma unroll
      for (int cd = 0; cd < NDIR; cd++) {
        sb += p.b[cd] * p.b[cd];
        sm += p.m[cd] * p.m[cd];
      }
      at(NE + 5, i) = p.rho;
      at(NE + 6, i) = p.e - 0.5 * (sb + sm * inv);
      at(NE + 7, i) = p.m[0] * inv;
    }
  }
  __syncthreads();
  const int c = c0 + t;
  const bool valid = c <= MhiK;
  const int k = t + 3;  // LDS index of the cell
  const bool fluxc = valid && c >= MloK + 2 && c <= MhiK - 2;
  const bool z1 = (c == MloK && !g.per_lo[0]) || (c == MhiK && !g.per_hi[0]);
  const double dxd = dx[valid ? c : MhiK];
  const double dxk = dxd, idx = 1.0 / dxd;
  double R[NE];
  {
    const double pTb0 = at(NE + 1, k), bgk = at(NE + 2, k), bgl = bgk;
    const double inv12 = idx / 12.0;
    double D1[NE], E1[NE];
#pragma unroll
    for (int n = 0; n < NE; n++) D1[n] = two_diff(at(n, k + 1), at(n, k - 1), E1[n]);
    const double dv1 = at(NE, k + 1) - at(NE, k - 1);
#pragma unroll
    for (int n = 0; n < NE; n++) {
      double e2;
      const double D2 = two_diff(at(n, k + 2), at(n, k - 2), e2);
      R[n] = ((D2 - 8.0 * D1[n]) + (e2 - 8.0 * E1[n])) * inv12;
    }
    double re = 0;
    const double Dv = -((at(NE, k + 2) - at(NE, k - 2)) - 8.0 * dv1) * inv12;
    re -= Dv * pTb0;
    re += bgk * bgl * Dv;
    R[1 + NDIR] += re;
#pragma unroll
    for (int n = 0; n < NE; n++) R[n] = fluxc ? R[n] : 0.0;
  }
  if (visc_on) {
    const double rtw0 = at(NE + 3, k - 1), rtw1 = at(NE + 3, k), rtw2 = at(NE + 3, k + 1);
#pragma unroll
    for (int pq = 0; pq < 3; pq++) {
#pragma clang fp contract(off)
      double q[7];
#pragma unroll
      for (int s2 = 0; s2 < 7; s2++) q[s2] = at(NE + 5 + pq, k - 3 + s2);
      const double ch = (pq == 2) ? 0.4 : 0.04;
      const double Kc = ctk * ch;
      double ratio0, ratio1;
      face_pair<double>(q, ratio0, ratio1);
      const double nuR = Kc * ratio1 * dxk;
      const double nuL = Kc * ratio0 * dxk;
      numax = valid ? fmax(numax, fmax(nuR, nuL)) : numax;
      const double qm = q[2], q0 = q[3], qp = q[4];
      const double gL = z1 ? 0.0 : (q0 - qm) * idx;
      const double gR = z1 ? 0.0 : (qp - q0) * idx;
      const double tL = nuL * gL, tR = nuR * gR;
      if (pq == 0) {
        R[0] += (tR - tL) * idx;
      } else if (pq == 1) {
        R[1 + NDIR] += (tR - tL) * idx;
      } else {
        const double rhoL = (rtw1 + rtw0) * 0.5, rhoR = (rtw2 + rtw1) * 0.5;
        const double m0 = at(NE + 4, k);
        const double mL = (m0 + at(NE + 4, k - 1)) * 0.5;
        const double mR = (at(NE + 4, k + 1) + m0) * 0.5;
        R[1] += (rhoR * tR - rhoL * tL) * idx;
        R[1 + NDIR] += (mR * tR - mL * tL) * idx;
      }
    }
  }
  if (valid) {
#pragma unroll
    for (int n = 0; n < NE; n++) Rout[n * fs + c] = R[n];
  }
}

// ---------------------------------------------------------------------
// kernel 2: gravity, derivatives of the cross-term factors, the register
// update out = base + qa*R (in RK4's last stage plus the reference's
// (w1 + 2*w2 + w3 - 4*w)/3, src/vac.t:450-454) and the cmax reductions of
// the result.
//   m_k += d_l(rho_t Tm_{k,l}),  e += d_l(m_k Tm_{k,l})
//   b_k -= d_l(Tb_{k,l}),        e -= d_l(b_k Tb_{k,l})
// with central differences along l, zero at a non-periodic edge of l, and
// T = 0 outside the mesh (nu is never set there, src/vacusr.viscosity.t:116-129).
// The register update runs in FP64 and is rounded once to S.

// V consecutive mesh cells along x from cell c of the grid g, of which the
// first nv are stored (the others, past the mesh, are evaluated on ghost
// cells and dropped); the max of the cmax values of the stored results goes
// to redv.  Each statement handles the V cells together, so that the loads
// of neighbouring cells (FP32 storage: 4 bytes apart) can be issued as one
// wider load; the arithmetic of every cell is that of V = 1.  RK4L: RK4's
// last stage (a.w1, a.w2 set; w3 is the stage input a.wct).
template <int NDIM, int NDIR, typename S, typename C, int V, bool RK4L>
__device__ __forceinline__ void final_cells(const FinalArgs& a, const GridDesc& g,
                                            const int (&c)[3], int nv,
                                            double (&redv)[2 * NDIM]) {
  constexpr int NE = 2 + 2 * NDIR;
  const long long fs = g.fs;
  const S* __restrict__ W = (const S*)a.wct;
  const S* __restrict__ RR = (const S*)a.R;
  const S* __restrict__ TT = (const S*)a.T;
  const double* __restrict__ B = a.bg;
  const long long ic0 = lin(g, c[0], c[1], c[2]);
  C R[V][NE];
#pragma unroll
  for (int n = 0; n < NE; n++)
#pragma unroll
    for (int j = 0; j < V; j++) R[j][n] = (C)RR[n * fs + ic0 + j];
  // gravity on the perturbed density (addsource_grav, src/vacusr.gravity.t:46-78)
  {
    // loads first, then selects (no branch between the loads)
    C rho[V], inv[V], md[V][NDIM];
#pragma unroll
    for (int j = 0; j < V; j++) rho[j] = (C)W[ic0 + j];
#pragma unroll
    for (int j = 0; j < V; j++) inv[j] = (C)recip<C>((double)W[ic0 + j] + B[fs + ic0 + j]);
#pragma unroll
    for (int d = 0; d < NDIM; d++)
#pragma unroll
      for (int j = 0; j < V; j++) md[j][d] = (C)W[(1 + d) * fs + ic0 + j];
#pragma unroll
    for (int j = 0; j < V; j++) {
      C rg = 0;
#pragma unroll
      for (int d = 0; d < NDIM; d++) {
        const bool on = a.grav_on[d] != 0;
        const C gd = (C)a.grav[d];
        R[j][1 + d] = on ? R[j][1 + d] + gd * rho[j] : R[j][1 + d];
        rg = on ? rg + gd * rho[j] * md[j][d] * inv[j] : rg;
      }
      R[j][1 + NDIR] += rg;
    }
  }
  if (a.visc_on && NDIM > 1) {
    // Branch-free, so that every load of the cell can be issued at once: the
    // neighbours c +- e_l of a mesh cell are in the grid; T outside the mesh
    // (never written) and the edge cells' terms are replaced by zero with
    // selects, never multiplied in.
#pragma unroll
    for (int l = 0; l < NDIM; l++) {
      if (l == NDIM - 1 && a.skip_last) continue;  // added by the line kernel of axis l
      const long long sl = g.st[l];
      bool edge[V], inP[V], inM[V];
      C h[V], rtp[V], rtm[V];
      long long ip[V], im[V];
#pragma unroll
      for (int j = 0; j < V; j++) {
        const int cl = c[l] + ((l == 0) ? j : 0);
        edge[j] = (cl == g.Mlo[l] && !g.per_lo[l]) || (cl == g.Mhi[l] && !g.per_hi[l]);
        h[j] = (C)(0.5 / a.dx[l][cl]);
        inP[j] = cl + 1 <= g.Mhi[l];
        inM[j] = cl - 1 >= g.Mlo[l];
        ip[j] = ic0 + j + sl;
        im[j] = ic0 + j - sl;
      }
#pragma unroll
      for (int j = 0; j < V; j++) rtp[j] = (C)((double)W[ip[j]] + B[fs + ip[j]]);
#pragma unroll
      for (int j = 0; j < V; j++) rtm[j] = (C)((double)W[im[j]] + B[fs + im[j]]);
#pragma unroll
      for (int k = 0; k < NDIM; k++) {
        if (k == l) continue;
        if (a.skip_xy && k + l == 1) continue;  // added by the plane tiles
        const long long xi = (long long)cross_index(NDIM, k, l) * fs;
        C Tmp0[V], Tmm0[V], Tbp0[V], Tbm0[V], mkp[V], mkm[V], bkp[V], bkm[V];
#pragma unroll
        for (int j = 0; j < V; j++) Tmp0[j] = (C)TT[xi + ip[j]];
#pragma unroll
        for (int j = 0; j < V; j++) Tmm0[j] = (C)TT[xi + im[j]];
#pragma unroll
        for (int j = 0; j < V; j++) Tbp0[j] = (C)TT[xi + fs + ip[j]];
#pragma unroll
        for (int j = 0; j < V; j++) Tbm0[j] = (C)TT[xi + fs + im[j]];
#pragma unroll
        for (int j = 0; j < V; j++) mkp[j] = (C)W[(1 + k) * fs + ip[j]];
#pragma unroll
        for (int j = 0; j < V; j++) mkm[j] = (C)W[(1 + k) * fs + im[j]];
#pragma unroll
        for (int j = 0; j < V; j++) bkp[j] = (C)W[(2 + NDIR + k) * fs + ip[j]];
#pragma unroll
        for (int j = 0; j < V; j++) bkm[j] = (C)W[(2 + NDIR + k) * fs + im[j]];
#pragma unroll
        for (int j = 0; j < V; j++) {
          const C Tmp = inP[j] ? Tmp0[j] : (C)0, Tmm = inM[j] ? Tmm0[j] : (C)0;
          const C Tbp = inP[j] ? Tbp0[j] : (C)0, Tbm = inM[j] ? Tbm0[j] : (C)0;
          const C dm = (rtp[j] * Tmp - rtm[j] * Tmm) * h[j];
          const C de1 = (mkp[j] * Tmp - mkm[j] * Tmm) * h[j];
          const C db = (Tbp - Tbm) * h[j];
          const C de2 = (Tbp * bkp[j] - Tbm * bkm[j]) * h[j];
          R[j][1 + k] += edge[j] ? (C)0 : dm;
          R[j][1 + NDIR] += edge[j] ? (C)0 : de1;
          R[j][2 + NDIR + k] -= edge[j] ? (C)0 : db;
          R[j][1 + NDIR] -= edge[j] ? (C)0 : de2;
        }
      }
    }
  }
  // register update and cmax of the (stored) result.  out may be base (the
  // in-place stages): each element is read by the thread that writes it,
  // before it writes it, so the two pointers carry no no-alias promise.
  S* out = (S*)a.out;
  const S* base = (const S*)a.base;
  S o[V][NE];
  // RK4's last stage: the reference's w1 = (w1 + 2*w2 + w3 - 4*w)/3 of the
  // three earlier stage registers and the step's base w, added to the update
  // (its left-to-right order kept, no contraction).  w3 is this stage's
  // input, so its values at the cell are the loads above where they exist.
  const S* __restrict__ r1 = (const S*)a.w1;
  const S* __restrict__ r2 = (const S*)a.w2;
  const S* __restrict__ r3 = W;
  auto rk4sum = [&](long long i, double b) {
#pragma clang fp contract(off)
    return ((((double)r1[i] + 2.0 * (double)r2[i]) + (double)r3[i]) - 4.0 * b) / 3.0;
  };
  if constexpr (V == 1) {
    // each field stored as soon as it is formed: measured faster at FP64
    // (1.46 against 1.48 ms per 256^3 launch) than all loads before the stores
    if constexpr (RK4L) {
#pragma unroll
      for (int n = 0; n < NE; n++) {
        const double b = (double)base[n * fs + ic0];
        const double u = b + a.qa * (double)R[0][n];
        o[0][n] = (S)(u + rk4sum(n * fs + ic0, b));
        out[n * fs + ic0] = o[0][n];
      }
    } else {
#pragma unroll
      for (int n = 0; n < NE; n++) {
        o[0][n] = (S)((double)base[n * fs + ic0] + a.qa * (double)R[0][n]);
        out[n * fs + ic0] = o[0][n];
      }
    }
  } else {
    if constexpr (RK4L) {
#pragma unroll
      for (int n = 0; n < NE; n++)
#pragma unroll
        for (int j = 0; j < V; j++) {
          const double b = (double)base[n * fs + ic0 + j];
          const double u = b + a.qa * (double)R[j][n];
          o[j][n] = (S)(u + rk4sum(n * fs + ic0 + j, b));
        }
    } else {
#pragma unroll
      for (int n = 0; n < NE; n++)
#pragma unroll
        for (int j = 0; j < V; j++)
          o[j][n] = (S)((double)base[n * fs + ic0 + j] + a.qa * (double)R[j][n]);
    }
    // all V cells stored in one branch (paired stores), the first nv in the other
    if (nv == V) {
#pragma unroll
      for (int n = 0; n < NE; n++)
#pragma unroll
        for (int j = 0; j < V; j++) out[n * fs + ic0 + j] = o[j][n];
    } else {
#pragma unroll
      for (int n = 0; n < NE; n++)
#pragma unroll
        for (int j = 0; j < V; j++)
          if (j < nv) out[n * fs + ic0 + j] = o[j][n];
    }
  }
#pragma unroll
  for (int j = 0; j < V; j++) {
    Pt<NDIR, double> po;
    po.rho = (double)o[j][0];
#pragma unroll
    for (int cd = 0; cd < NDIR; cd++) {
      po.m[cd] = (double)o[j][1 + cd];
      po.b[cd] = (double)o[j][2 + NDIR + cd];
      po.bg[cd] = B[(2 + cd) * fs + ic0 + j];
    }
    po.e = (double)o[j][1 + NDIR];
    po.eb = B[ic0 + j];
    po.rhob = B[fs + ic0 + j];
    double cm[NDIM];
    cmax_all<NDIM, NDIR>(po, a.gamma, cm);
#pragma unroll
    for (int d = 0; d < NDIM; d++) {
      const double r0 = cm[d], r1 = cm[d] / a.dx[d][c[d] + ((d == 0) ? j : 0)];
      if (j == 0) {
        redv[d] = r0;
        redv[NDIM + d] = r1;
      } else if (j < nv) {
        redv[d] = fmax(redv[d], r0);
        redv[NDIM + d] = fmax(redv[NDIM + d], r1);
      }
    }
  }
}

// One mesh cell c of the grid g; the cmax values of the result go to redv.
// RK4L: RK4's last stage (a.w1, a.w2 set; w3 is the stage input a.wct).
template <int NDIM, int NDIR, typename S, typename C, bool RK4L>
__device__ __forceinline__ void final_cell(const FinalArgs& a, const GridDesc& g, const int (&c)[3],
                                           double (&redv)[2 * NDIM]) {
  final_cells<NDIM, NDIR, S, C, 1, RK4L>(a, g, c, 1, redv);
}

// ---------------------------------------------------------------------
// kernel 4: ghost-cell fill.  The reference fills boundary regions one
// after another (iB = 1..2*ndim), each spanning the whole grid along the
// other axes, so a cell that is ghost along several axes ends up with the
// value written by its highest such axis, taken from a cell that was
// filled earlier.  Following that chain backwards reaches a mesh cell (or
// the stored value of a 'fixed' region), which is what every thread does
// here; no thread reads a ghost cell, so the kernel is race-free.

// Value of ghost cell c of variable var after the reference's sequential
// region fills (src/vacgrid.t:195-512): the region of the highest axis along
// which c is a ghost writes last, and its sources lie in the mesh along that
// axis, so they are evaluated the same way on the lower axes.  D is the
// highest axis still to resolve.
template <typename S, int D>
__device__ double ghost_eval(const GhostArgs& a, int var, int c0, int c1, int c2) {
#pragma clang fp contract(off)
  const GridDesc& g = a.g;
  const long long vo = (long long)var * g.fs;
  if constexpr (D < 0) {
    return (double)((const S*)a.f)[vo + lin(g, c0, c1, c2)];
  } else {
    int c[3] = {c0, c1, c2};
    if (D >= g.ndim || (c[D] >= g.Mlo[D] && c[D] <= g.Mhi[D]))
      return ghost_eval<S, D - 1>(a, var, c0, c1, c2);
    const S* f = (const S*)a.f;
    const S* snap = (const S*)a.snap;
    const long long i = vo + lin(g, c[0], c[1], c[2]);
    const int side = (c[D] > g.Mhi[D]) ? 1 : 0;
    const int code = a.bc[var * 6 + 2 * D + side];
    const int N = g.Mhi[D] - g.Mlo[D] + 1;
    const int ixe = side ? g.Mhi[D] : g.Mlo[D];  // edge cell
    auto at = [&](int cd) {
      int e[3] = {c[0], c[1], c[2]};
      e[D] = cd;
      return ghost_eval<S, D - 1>(a, var, e[0], e[1], e[2]);
    };
    switch (code) {
      case BC_PERIODIC:
        return at(c[D] + (side ? -N : N));
      case BC_SYMM:
      case BC_SYMM0:
        return at(side ? (2 * g.Mhi[D] + 1 - c[D]) : (2 * g.Mlo[D] - 1 - c[D]));
      case BC_ASYMM:
        return -at(side ? (2 * g.Mhi[D] + 1 - c[D]) : (2 * g.Mlo[D] - 1 - c[D]));
      case BC_CONTCD4:  // mirror about the edge cell
        return at(2 * ixe - c[D]);
      case BC_FIXED:
      case BC_FIXED1:
      case BC_GRAD1:
        if (snap) {
          if (code == BC_GRAD1) return (double)snap[i] + at(ixe);
          return (double)snap[i];
        }
        if (a.keep_fixed) return (double)f[i];  // fullgridini: file values
        if (code == BC_FIXED) return at(ixe);
        [[fallthrough]];
      case BC_CONT1: {
        const int k = abs(c[D] - ixe);
        const int ixf = side ? ixe - 1 : ixe + 1;
        return (double)(k + 1) * at(ixe) - (double)k * at(ixf);
      }
      default:  // cont
        return at(ixe);
    }
  }
}

// The common path of ghost_eval, iterative: single-source fills move the
// cell into the mesh axis by axis, keeping result = sgn * value + add; the
// two-source extrapolations (cont1, and fixed1/grad1 before their values
// are stored) hand the rest to ghost_eval, compiled only into the GENERAL
// kernel (the launcher picks it when such a fill occurs).
template <typename S, bool GENERAL>
__device__ __forceinline__ S ghost_value(const GhostArgs& a, int var, int c0, int c1, int c2) {
  const GridDesc& g = a.g;
  const S* f = (const S*)a.f;
  const S* snap = (const S*)a.snap;
  const long long vo = (long long)var * g.fs;
  int c[3] = {c0, c1, c2};
  double sgn = 1.0, add = 0.0;
  for (int iter = 0; iter < 3; iter++) {
    int d = -1;
#pragma unroll
    for (int dd = 2; dd >= 0; dd--) {
      if (d < 0 && dd < g.ndim && (c[dd] < g.Mlo[dd] || c[dd] > g.Mhi[dd])) d = dd;
    }
    const long long i = vo + lin(g, c[0], c[1], c[2]);
    if (d < 0) break;
    const int side = (c[d] > g.Mhi[d]) ? 1 : 0;
    const int code = a.bc[var * 6 + 2 * d + side];
    const int N = g.Mhi[d] - g.Mlo[d] + 1;
    const int ixe = side ? g.Mhi[d] : g.Mlo[d];
    const bool stored = code == BC_FIXED || code == BC_FIXED1 || code == BC_GRAD1;
    if (code == BC_PERIODIC) {
      c[d] += side ? -N : N;
    } else if (stored && snap) {
      if (code != BC_GRAD1) return (S)(sgn * (double)snap[i] + add);
      add += sgn * (double)snap[i];
      c[d] = ixe;
    } else if (stored && a.keep_fixed) {
      return (S)(sgn * (double)f[i] + add);
    } else if (code == BC_CONT || code == BC_FIXED) {
      c[d] = ixe;
    } else if (code == BC_CONTCD4) {
      c[d] = 2 * ixe - c[d];
    } else if (code == BC_SYMM || code == BC_SYMM0 || code == BC_ASYMM) {
      c[d] = side ? (2 * g.Mhi[d] + 1 - c[d]) : (2 * g.Mlo[d] - 1 - c[d]);
      if (code == BC_ASYMM) sgn = -sgn;
    } else {  // two-source extrapolation: the general evaluator from here
      if constexpr (GENERAL) {
        double v;
        if (d == 2) v = ghost_eval<S, 2>(a, var, c[0], c[1], c[2]);
        else if (d == 1) v = ghost_eval<S, 1>(a, var, c[0], c[1], c[2]);
        else v = ghost_eval<S, 0>(a, var, c[0], c[1], c[2]);
        return (S)(sgn * v + add);
      } else {
        return (S)0;  // not reached: the launcher selects the GENERAL kernel
      }
    }
  }
  return (S)(sgn * (double)f[vo + lin(g, c[0], c[1], c[2])] + add);
}

// Stored value of ghost cell c for the boundaries that keep one (fixed,
// fixed1: the value; grad1: the difference to the edge cell of its axis,
// src/vacgrid.t:476-507), from the filled field; 0 for other types.
template <typename S>
__device__ __forceinline__ S ghost_stored(const GhostArgs& a, int var, int c0, int c1, int c2) {
  const GridDesc& g = a.g;
  const S* f = (const S*)a.f;
  const long long vo = (long long)var * g.fs;
  int c[3] = {c0, c1, c2};
  int d = -1;
#pragma unroll
  for (int dd = 2; dd >= 0; dd--)
    if (d < 0 && dd < g.ndim && (c[dd] < g.Mlo[dd] || c[dd] > g.Mhi[dd])) d = dd;
  if (d < 0) return (S)0;
  const int side = (c[d] > g.Mhi[d]) ? 1 : 0;
  const int code = a.bc[var * 6 + 2 * d + side];
  const S v = f[vo + lin(g, c[0], c[1], c[2])];
  if (code == BC_FIXED || code == BC_FIXED1) return v;
  if (code == BC_GRAD1) {
    c[d] = side ? g.Mhi[d] : g.Mlo[d];
    return (S)((double)v - (double)f[vo + lin(g, c[0], c[1], c[2])]);
  }
  return (S)0;
}

}  // namespace sac
