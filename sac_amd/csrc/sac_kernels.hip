// sac_kernels.hip -- HIP kernels (gfx950) of the SAC explicit
// perturbation-MHD stage: face ratios of the hyperviscosity coefficient,
// the fused stage update (flux divergence, background energy terms,
// gravity, hyperdiffusion, Runge-Kutta register update and the cmax /
// maxviscoef reductions of the result), ghost-cell fill and reductions.
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
// point state

template <int NDIR>
struct Pt {
  double rho, m[NDIR], e, b[NDIR];  // perturbation (evolving)
  double eb, rhob, bg[NDIR];        // static background
};

template <int NDIR>
__device__ __forceinline__ void load_pt(const double* __restrict__ W, const double* __restrict__ B,
                                        long long fs, long long i, Pt<NDIR>& p) {
  p.rho = W[i];
#pragma unroll
  for (int c = 0; c < NDIR; c++) p.m[c] = W[(1 + c) * fs + i];
  p.e = W[(1 + NDIR) * fs + i];
#pragma unroll
  for (int c = 0; c < NDIR; c++) p.b[c] = W[(2 + NDIR + c) * fs + i];
  p.eb = B[i];
  p.rhob = B[fs + i];
#pragma unroll
  for (int c = 0; c < NDIR; c++) p.bg[c] = B[(2 + c) * fs + i];
}

// perturbation total pressure p_T' (getptotal, src/vacphys0.t.mhd:186-205)
template <int NDIR>
__device__ __forceinline__ double ptotal(const Pt<NDIR>& p, double gamma, double inv) {
  double sbbg = 0.0, sb2 = 0.0, sm2 = 0.0;
#pragma unroll
  for (int c = 0; c < NDIR; c++) {
    sbbg += p.b[c] * p.bg[c];
    sb2 += p.b[c] * p.b[c];
    sm2 += p.m[c] * p.m[c];
  }
  return (gamma - 1.0) * (p.e - 0.5 * sm2 * inv) - (gamma - 2.0) * (sbbg + 0.5 * sb2);
}

// background total pressure (getptotal_bg, src/vacphys0.t.mhd:208-224)
template <int NDIR>
__device__ __forceinline__ double ptotal_bg(const Pt<NDIR>& p, double gamma) {
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < NDIR; c++) s += p.bg[c] * p.bg[c];
  return (gamma - 1.0) * p.eb - 0.5 * (gamma - 2.0) * s;
}

// full flux (non-transport part of getflux + transport v_d*w) along axis d
template <int NDIR>
__device__ __forceinline__ void flux_pt(const Pt<NDIR>& p, int d, double gamma,
                                        double (&F)[2 + 2 * NDIR]) {
  const double rt = p.rho + p.rhob;
  const double inv = 1.0 / rt;
  const double vd = p.m[d] * inv;
  F[0] = p.rhob * p.m[d] * inv + vd * p.rho;
  const double pT = ptotal<NDIR>(p, gamma, inv);
  double sbgm = 0.0, sbm = 0.0;
#pragma unroll
  for (int c = 0; c < NDIR; c++) {
    sbgm += p.bg[c] * p.m[c];
    sbm += p.b[c] * p.m[c];
  }
#pragma unroll
  for (int c = 0; c < NDIR; c++) {
    double f = -(p.b[c] * p.bg[d] + p.b[d] * p.bg[c]) - p.b[c] * p.b[d];
    if (c == d) f = pT + f;
    F[1 + c] = f + vd * p.m[c];
  }
  F[1 + NDIR] = (p.m[d] * pT - p.b[d] * sbgm - p.bg[d] * sbm) * inv + p.eb * p.m[d] * inv -
                p.b[d] * sbm * inv + vd * p.e;
#pragma unroll
  for (int c = 0; c < NDIR; c++) {
    F[2 + NDIR + c] = (c == d) ? 0.0
                               : (-p.m[c] * inv * (p.b[d] + p.bg[d]) + p.m[d] * inv * p.bg[c] +
                                  vd * p.b[c]);
  }
}

// fast-mode speed + |v_d| (getcmax, src/vacphys0.t.mhd:77-112)
template <int NDIR>
__device__ __forceinline__ double cmax_pt(const Pt<NDIR>& p, int d, double gamma) {
  const double rt = p.rho + p.rhob;
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
  double pth = 0.5 * sm2 / rt;
  pth = pth + 0.5 * sb2 + sbbg;
  pth = (gamma - 1.0) * (p.e - pth);
  const double cs2 = gamma * (pth + (gamma - 1.0) * (p.eb - 0.5 * sbg2)) / rt;
  const double cf2 = sbt2 / rt + cs2;
  const double bd = p.b[d] + p.bg[d];
  return sqrt(0.5 * (cf2 + sqrt(cf2 * cf2 - 4.0 * cs2 * (bd * bd) / rt))) + fabs(p.m[d] / rt);
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

// ---------------------------------------------------------------------
// setnu variables at one point (q of setnu for every variable)

template <int NDIR>
struct Qv {
  double rho, qe, rt, inv, v[3], m[3], b[3];
};

template <int NDIM, int NDIR>
__device__ __forceinline__ void load_qv(const double* __restrict__ W, const double* __restrict__ B,
                                        long long fs, long long i, Qv<NDIR>& q) {
  q.rho = W[i];
  q.rt = q.rho + B[fs + i];
  q.inv = 1.0 / q.rt;
  double sb = 0.0, sm = 0.0;
#pragma unroll
  for (int c = 0; c < NDIR; c++) {
    const double mc = W[(1 + c) * fs + i];
    const double bc = W[(2 + NDIR + c) * fs + i];
    sb += bc * bc;
    sm += mc * mc;
    if (c < NDIM) {
      q.m[c] = mc;
      q.v[c] = mc * q.inv;
      q.b[c] = bc;
    }
  }
  q.qe = W[(1 + NDIR) * fs + i] - 0.5 * (sb + sm * q.inv);
}

// setnu ratio md3/md1 at the face between q[j] and q[j+1] of a 6-value
// window q[j-2..j+3] (src/vacusr.viscosity.t:413-427)
__device__ __forceinline__ double face_ratio(const double* q) {
  double d1[5], d3[3];
#pragma unroll
  for (int j = 0; j < 5; j++) d1[j] = fabs(q[j + 1] - q[j]);
#pragma unroll
  for (int j = 0; j < 3; j++) d3[j] = fabs(3.0 * (q[j + 2] - q[j + 1]) - (q[j + 3] - q[j]));
  const double md3 = fmax(fmax(d3[0], d3[1]), d3[2]);
  const double md1 = fmax(fmax(fmax(d1[0], d1[1]), fmax(d1[2], d1[3])), d1[4]);
  return (md1 > 0.0) ? md3 / md1 : 0.0;
}

__device__ __forceinline__ int cross_index(int ndim, int k, int l) {
  return 2 * (k * (ndim - 1) + (l < k ? l : l - 1));
}

// ---------------------------------------------------------------------
// kernel 1: the work of one axis K for every mesh cell, all of which lies
// on the K-line through the cell:
//   - 4th-order flux divergence along K (centdiff4, src/vaccd.t:35-70)
//     and the background energy terms with derivatives along K,
//   - setnu (src/vacusr.viscosity.t:224-455) for the 2*ndim+1 variables of
//     axis K from the 7-point stencil (mirror layer beyond the grid),
//   - the hyperdiffusion terms along K of addsource_visc (rho, thermal
//     energy, momentum i==k branch, field m==k branch),
//   - the cross-term factors T (diffusion along K) that kernel 2
//     differentiates along the other axes:
//       Tm_{K,l} = d_K v_l * (nuL+nuR)_{m_l}/4,  Tb_{K,l} = d_K b_l * (nuL+nuR)_{b_l}/2.
// The contribution is added to the rate accumulator R (written on K=0).

// Workgroups are dealt to the 8 XCDs of the device in turn; renumbering them
// so that each XCD takes one contiguous range of cells keeps the neighbour
// rows/planes of a line stencil in the same XCD's L2.  The grid is a multiple
// of 8 blocks, so the renumbering is a bijection.
__device__ __forceinline__ long long xcd_block(int xcd) {
  const unsigned nb = gridDim.x;
  if (xcd <= 1 || (nb % (unsigned)xcd) != 0) return blockIdx.x;
  const unsigned per = nb / (unsigned)xcd;
  return (long long)(blockIdx.x % (unsigned)xcd) * per + blockIdx.x / (unsigned)xcd;
}

template <int NDIM, int NDIR, int K>
__global__ void __launch_bounds__(256) k_axis(AxisArgs a) {
  constexpr int NE = 2 + 2 * NDIR;
  constexpr int NP = 2 * NDIM + 1;
  const GridDesc& g = a.g;
  const long long t = xcd_block(a.xcd) * blockDim.x + threadIdx.x;
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = g.Mlo[d];
    ext[d] = g.Mhi[d] - g.Mlo[d] + 1;
  }
  const bool active = t < (long long)ext[0] * ext[1] * ext[2];
  double numax = 0.0;
  if (active) {
    int c[3];
    decode_box(t, lo, ext, c);
    const long long fs = g.fs;
    const double* __restrict__ W = a.wct;
    const double* __restrict__ B = a.bg;
    const double gamma = a.gamma;
    const long long ic = lin(g, c[0], c[1], c[2]);
    const long long sk = g.st[K];
    const double dxk = a.dx[K][c[K]];
    const double idx = 1.0 / dxk;
    double R[NE];
#pragma unroll
    for (int v = 0; v < NE; v++) R[v] = 0.0;

    // ---- flux divergence along K (zero on two layers at each side) ----
    if (!((c[K] <= g.Mlo[K] + 1) || (c[K] >= g.Mhi[K] - 1))) {
      double bgk, bgl[NDIM], pTb0;
      {
        Pt<NDIR> p0;
        load_pt<NDIR>(W, B, fs, ic, p0);
        pTb0 = ptotal_bg<NDIR>(p0, gamma);
        bgk = p0.bg[K];
#pragma unroll
        for (int l = 0; l < NDIM; l++) bgl[l] = p0.bg[l];
      }
      const double inv12 = idx / 12.0;
      double Fm[NE], D1[NE], vm[NDIM], dv1[NDIM];
      SCHED_FENCE();
      {
        Pt<NDIR> p;
        load_pt<NDIR>(W, B, fs, ic - sk, p);
        flux_pt<NDIR>(p, K, gamma, Fm);
        const double inv = 1.0 / (p.rho + p.rhob);
#pragma unroll
        for (int l = 0; l < NDIM; l++) vm[l] = p.m[l] * inv;
      }
      SCHED_FENCE();
      {
        Pt<NDIR> p;
        double F[NE];
        load_pt<NDIR>(W, B, fs, ic + sk, p);
        flux_pt<NDIR>(p, K, gamma, F);
        const double inv = 1.0 / (p.rho + p.rhob);
#pragma unroll
        for (int v = 0; v < NE; v++) D1[v] = F[v] - Fm[v];
#pragma unroll
        for (int l = 0; l < NDIM; l++) dv1[l] = p.m[l] * inv - vm[l];
      }
      SCHED_FENCE();
      {
        Pt<NDIR> p;
        load_pt<NDIR>(W, B, fs, ic - 2 * sk, p);
        flux_pt<NDIR>(p, K, gamma, Fm);
        const double inv = 1.0 / (p.rho + p.rhob);
#pragma unroll
        for (int l = 0; l < NDIM; l++) vm[l] = p.m[l] * inv;
      }
      SCHED_FENCE();
      {
        Pt<NDIR> p;
        double F[NE];
        load_pt<NDIR>(W, B, fs, ic + 2 * sk, p);
        flux_pt<NDIR>(p, K, gamma, F);
        const double inv = 1.0 / (p.rho + p.rhob);
        // -gradient4 = ((F2-F-2) - 8 (F1-F-1)) / (12 dx)
#pragma unroll
        for (int v = 0; v < NE; v++) R[v] += ((F[v] - Fm[v]) - 8.0 * D1[v]) * inv12;
        double re = 0.0;
#pragma unroll
        for (int l = 0; l < NDIM; l++) {
          const double Dv = -((p.m[l] * inv - vm[l]) - 8.0 * dv1[l]) * inv12;
          if (l == K) re -= Dv * pTb0;
          re += bgk * bgl[l] * Dv;
        }
        R[1 + NDIR] += re;
      }
    }
    SCHED_FENCE();

    // ---- hyperdiffusion along K ----
    if (a.visc_on) {
      const double ctk = red_get(a.ctot_src, K);
      // rho_t and 1/rho_t at offsets -3..3 (mirror beyond the grid)
      long long id7[7];
      double inv7[7];
#pragma unroll
      for (int s = 0; s < 7; s++) {
        int j = c[K] + s - 3;
        if (j < 0) j = g.Mlo[K] + 3;
        if (j > g.G[K] - 1) j = g.Mhi[K] - 3;
        id7[s] = ic + (long long)(j - c[K]) * sk;
        inv7[s] = 1.0 / (W[id7[s]] + B[fs + id7[s]]);
      }
      const bool z1 = (c[K] == g.Mlo[K] && !g.per_lo[K]) || (c[K] == g.Mhi[K] && !g.per_hi[K]);
      const double rtm = W[ic - sk] + B[fs + ic - sk];
      const double rt0 = W[ic] + B[fs + ic];
      const double rtp = W[ic + sk] + B[fs + ic + sk];
      // p = 0: rho' ; p = 1: thermal energy ; p = 2+l: v_l ; then b_l (l != K)
#pragma unroll
      for (int p = 0; p < NP; p++) {
        double q[7];
        int lb = -1;
        if (p >= 2 + NDIM) {
          int n = p - 2 - NDIM;
          lb = (n < K) ? n : n + 1;
        }
#pragma unroll
        for (int s = 0; s < 7; s++) {
          const long long i = id7[s];
          if (p == 0) {
            q[s] = W[i];
          } else if (p == 1) {
            double sb = 0.0, sm = 0.0;
#pragma unroll
            for (int cd = 0; cd < NDIR; cd++) {
              const double mc = W[(1 + cd) * fs + i];
              const double bc = W[(2 + NDIR + cd) * fs + i];
              sb += bc * bc;
              sm += mc * mc;
            }
            q[s] = W[(1 + NDIR) * fs + i] - 0.5 * (sb + sm * inv7[s]);
          } else if (p < 2 + NDIM) {
            q[s] = W[(p - 1) * fs + i] * inv7[s];
          } else {
            q[s] = W[(2 + NDIR + lb) * fs + i];
          }
        }
        const double ch = (p >= 2 && p < 2 + NDIM) ? 0.4 : 0.04;
        const double Kc = ctk * ch;
        const double nuR = Kc * face_ratio(q + 1) * dxk;
        const double nuL = Kc * face_ratio(q) * dxk;
        numax = fmax(numax, fmax(nuR, nuL));
        const double gL = z1 ? 0.0 : (q[3] - q[2]) * idx;
        const double gR = z1 ? 0.0 : (q[4] - q[3]) * idx;
        const double tL = nuL * gL, tR = nuR * gR;
        if (p == 0) {
          R[0] += (tR - tL) * idx;
        } else if (p == 1) {
          R[1 + NDIR] += (tR - tL) * idx;
        } else if (p < 2 + NDIM) {
          const int l = p - 2;
          const double rhoL = (rt0 + rtm) * 0.5, rhoR = (rtp + rt0) * 0.5;
          const double m0 = W[(1 + l) * fs + ic];
          const double mL = (m0 + W[(1 + l) * fs + ic - sk]) * 0.5;
          const double mR = (W[(1 + l) * fs + ic + sk] + m0) * 0.5;
          const double mult = (l == K) ? 1.0 : 0.5;  // the i==k branch runs twice for l==k
          R[1 + l] += mult * (rhoR * tR - rhoL * tL) * idx;
          R[1 + NDIR] += mult * (mR * tR - mL * tL) * idx;
          if (l != K) {
            const double gC = z1 ? 0.0 : (q[4] - q[2]) * idx * 0.5;
            a.T[(long long)cross_index(NDIM, K, l) * fs + ic] = gC * (nuL + nuR) * 0.25;
          }
        } else {
          const double BL = (q[3] + q[2]) * 0.5, BR = (q[4] + q[3]) * 0.5;
          R[2 + NDIR + lb] += (tR - tL) * idx;
          R[1 + NDIR] += (tR * BR - tL * BL) * idx;
          const double gC = z1 ? 0.0 : (q[4] - q[2]) * idx * 0.5;
          a.T[(long long)(cross_index(NDIM, K, lb) + 1) * fs + ic] = gC * (nuL + nuR) * 0.5;
        }
        SCHED_FENCE();
      }
    }
    // ---- accumulate ----
    double* __restrict__ Racc = a.R;
    if (K == 0) {
#pragma unroll
      for (int v = 0; v < NE; v++) Racc[v * fs + ic] = R[v];
    } else {
#pragma unroll
      for (int v = 0; v < NE; v++) Racc[v * fs + ic] += R[v];
    }
  }
  if (a.visc_on) {
    double r2[1] = {numax};
    block_max_atomic<1>(r2, a.red_visc);
  }
}

// ---------------------------------------------------------------------
// kernel 2: gravity, derivatives of the cross-term factors, the register
// update out = base + qa*R (acc_out = acc_in + qb*R) and the cmax
// reductions of the result.
//   m_k += d_l(rho_t Tm_{k,l}),  e += d_l(m_k Tm_{k,l})
//   b_k -= d_l(Tb_{k,l}),        e -= d_l(b_k Tb_{k,l})
// with central differences along l, zero at a non-periodic edge of l, and
// T = 0 outside the mesh (nu is never set there, src/vacusr.viscosity.t:116-129).

template <int NDIM, int NDIR>
__global__ void __launch_bounds__(256) k_final(FinalArgs a) {
  constexpr int NE = 2 + 2 * NDIR;
  const GridDesc& g = a.g;
  const long long t = xcd_block(a.xcd) * blockDim.x + threadIdx.x;
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = g.Mlo[d];
    ext[d] = g.Mhi[d] - g.Mlo[d] + 1;
  }
  const bool active = t < (long long)ext[0] * ext[1] * ext[2];
  double redv[2 * NDIM];
#pragma unroll
  for (int k = 0; k < 2 * NDIM; k++) redv[k] = 0.0;
  if (active) {
    int c[3];
    decode_box(t, lo, ext, c);
    const long long fs = g.fs;
    const double* __restrict__ W = a.wct;
    const double* __restrict__ B = a.bg;
    const long long ic = lin(g, c[0], c[1], c[2]);
    double R[NE];
#pragma unroll
    for (int v = 0; v < NE; v++) R[v] = a.R[v * fs + ic];
    // gravity on the perturbed density (addsource_grav, src/vacusr.gravity.t:46-78)
    {
      const double rho = W[ic];
      const double inv = 1.0 / (rho + B[fs + ic]);
      double rg = 0.0;
#pragma unroll
      for (int d = 0; d < NDIM; d++) {
        if (a.grav_on[d]) {
          R[1 + d] += a.grav[d] * rho;
          rg += a.grav[d] * rho * W[(1 + d) * fs + ic] * inv;
        }
      }
      R[1 + NDIR] += rg;
    }
    if (a.visc_on && NDIM > 1) {
#pragma unroll
      for (int l = 0; l < NDIM; l++) {
        if ((c[l] == g.Mlo[l] && !g.per_lo[l]) || (c[l] == g.Mhi[l] && !g.per_hi[l])) continue;
        const long long sl = g.st[l];
        const double h = 0.5 / a.dx[l][c[l]];
        const bool inP = c[l] + 1 <= g.Mhi[l], inM = c[l] - 1 >= g.Mlo[l];
        const long long ip = ic + sl, im = ic - sl;
        const double rtp = W[ip] + B[fs + ip], rtm = W[im] + B[fs + im];
#pragma unroll
        for (int k = 0; k < NDIM; k++) {
          if (k == l) continue;
          const long long xi = (long long)cross_index(NDIM, k, l) * fs;
          const double Tmp = inP ? a.T[xi + ip] : 0.0, Tmm = inM ? a.T[xi + im] : 0.0;
          const double Tbp = inP ? a.T[xi + fs + ip] : 0.0, Tbm = inM ? a.T[xi + fs + im] : 0.0;
          const double mkp = W[(1 + k) * fs + ip], mkm = W[(1 + k) * fs + im];
          const double bkp = W[(2 + NDIR + k) * fs + ip], bkm = W[(2 + NDIR + k) * fs + im];
          R[1 + k] += (rtp * Tmp - rtm * Tmm) * h;
          R[1 + NDIR] += (mkp * Tmp - mkm * Tmm) * h;
          R[2 + NDIR + k] -= (Tbp - Tbm) * h;
          R[1 + NDIR] -= (Tbp * bkp - Tbm * bkm) * h;
        }
      }
    }
    // register update and cmax of the result
    Pt<NDIR> po;
    {
      double vals[NE];
#pragma unroll
      for (int v = 0; v < NE; v++) {
        vals[v] = a.base[v * fs + ic] + a.qa * R[v];
        a.out[v * fs + ic] = vals[v];
      }
      if (a.acc_out) {
#pragma unroll
        for (int v = 0; v < NE; v++) a.acc_out[v * fs + ic] = a.acc_in[v * fs + ic] + a.qb * R[v];
      }
      po.rho = vals[0];
#pragma unroll
      for (int cd = 0; cd < NDIR; cd++) {
        po.m[cd] = vals[1 + cd];
        po.b[cd] = vals[2 + NDIR + cd];
        po.bg[cd] = B[(2 + cd) * fs + ic];
      }
      po.e = vals[1 + NDIR];
      po.eb = B[ic];
      po.rhob = B[fs + ic];
    }
#pragma unroll
    for (int d = 0; d < NDIM; d++) {
      const double cm = cmax_pt<NDIR>(po, d, a.gamma);
      redv[d] = cm;
      redv[NDIM + d] = cm / a.dx[d][c[d]];
    }
  }
  block_max_atomic<2 * NDIM>(redv, a.red_out);
}

// ---------------------------------------------------------------------
// kernel 3: cmax reductions of a register (getdt_courant / setnu c_tot)

template <int NDIM, int NDIR>
__global__ void __launch_bounds__(256) k_cmax(CmaxArgs a) {
  const GridDesc& g = a.g;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = g.Mlo[d];
    ext[d] = g.Mhi[d] - g.Mlo[d] + 1;
  }
  double r[2 * NDIM];
#pragma unroll
  for (int k = 0; k < 2 * NDIM; k++) r[k] = 0.0;
  if (t < (long long)ext[0] * ext[1] * ext[2]) {
    int c[3];
    decode_box(t, lo, ext, c);
    Pt<NDIR> p;
    load_pt<NDIR>(a.w, a.bg, g.fs, lin(g, c[0], c[1], c[2]), p);
#pragma unroll
    for (int d = 0; d < NDIM; d++) {
      const double cm = cmax_pt<NDIR>(p, d, a.gamma);
      r[d] = cm;
      r[NDIM + d] = cm / a.dx[d][c[d]];
    }
  }
  block_max_atomic<2 * NDIM>(r, a.red_out);
}

// ---------------------------------------------------------------------
// kernel 4: ghost-cell fill.  The reference fills boundary regions one
// after another (iB = 1..2*ndim), each spanning the whole grid along the
// other axes, so a cell that is ghost along several axes ends up with the
// value written by its highest such axis, taken from a cell that was
// filled earlier.  Following that chain backwards reaches a mesh cell (or
// the stored value of a 'fixed' region), which is what every thread does
// here; no thread reads a ghost cell, so the kernel is race-free.

__device__ __forceinline__ double ghost_value(const GhostArgs& a, int var, int c0, int c1, int c2) {
  const GridDesc& g = a.g;
  int c[3] = {c0, c1, c2};
  double sign = 1.0;
  for (int iter = 0; iter < 8; iter++) {
    int d = -1;
#pragma unroll
    for (int dd = 2; dd >= 0; dd--) {
      if (d < 0 && dd < g.ndim && (c[dd] < g.Mlo[dd] || c[dd] > g.Mhi[dd])) d = dd;
    }
    const long long i = lin(g, c[0], c[1], c[2]);
    if (d < 0) return sign * a.f[(long long)var * g.fs + i];
    const int side = (c[d] > g.Mhi[d]) ? 1 : 0;
    const int code = a.bc[var * 6 + 2 * d + side];
    const int N = g.Mhi[d] - g.Mlo[d] + 1;
    if (code == BC_PERIODIC) {
      c[d] += side ? -N : N;
    } else if (code == BC_FIXED && a.snap) {
      return sign * a.snap[(long long)var * g.fs + i];
    } else if (code == BC_FIXED && a.keep_fixed) {
      return sign * a.f[(long long)var * g.fs + i];
    } else if (code == BC_CONT || code == BC_FIXED) {
      c[d] = side ? g.Mhi[d] : g.Mlo[d];
    } else {  // symm, symm0, asymm: reflection about the boundary face
      c[d] = side ? (2 * g.Mhi[d] + 1 - c[d]) : (2 * g.Mlo[d] - 1 - c[d]);
      if (code == BC_ASYMM) sign = -sign;
    }
  }
  return 0.0;
}

__global__ void __launch_bounds__(256) k_ghost(GhostArgs a) {
  const GridDesc& g = a.g;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  // ghost shell = part A (ghost along axis 2 [0-based], any i0,i1)
  //             + part B (mesh along 2, ghost along 1, any i0)
  //             + part C (mesh along 2 and 1, ghost along 0)
  const int G0 = g.G[0], G1 = g.G[1], G2 = g.G[2];
  const int N0 = g.Mhi[0] - g.Mlo[0] + 1;
  const int N1 = (g.ndim > 1) ? g.Mhi[1] - g.Mlo[1] + 1 : 1;
  const int N2 = (g.ndim > 2) ? g.Mhi[2] - g.Mlo[2] + 1 : 1;
  const int gh2 = (g.ndim > 2) ? G2 - N2 : 0;
  const int gh1 = (g.ndim > 1) ? G1 - N1 : 0;
  const int gh0 = G0 - N0;
  const long long nA = (long long)gh2 * G1 * G0;
  const long long nB = (long long)N2 * gh1 * G0;
  const long long nC = (long long)N2 * N1 * gh0;
  int c[3];
  if (t < nA) {
    c[0] = (int)(t % G0);
    long long r = t / G0;
    c[1] = (int)(r % G1);
    int j = (int)(r / G1);
    c[2] = (j < g.Mlo[2]) ? j : j + N2;
  } else if (t < nA + nB) {
    long long u = t - nA;
    c[0] = (int)(u % G0);
    long long r = u / G0;
    int j = (int)(r % gh1);
    c[1] = (j < g.Mlo[1]) ? j : j + N1;
    c[2] = g.Mlo[2] + (int)(r / gh1);
  } else if (t < nA + nB + nC) {
    long long u = t - nA - nB;
    int j = (int)(u % gh0);
    c[0] = (j < g.Mlo[0]) ? j : j + N0;
    long long r = u / gh0;
    c[1] = g.Mlo[1] + (int)(r % N1);
    c[2] = g.Mlo[2] + (int)(r / N1);
  } else {
    return;
  }
  const long long io = lin(g, c[0], c[1], c[2]);
  for (int v = 0; v < a.nvar; v++) {
    const double val = ghost_value(a, v, c[0], c[1], c[2]);
    a.f[(long long)v * g.fs + io] = val;
  }
}

// ---------------------------------------------------------------------
// small helpers

__global__ void k_zero_u64(unsigned long long* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// w3 <- (w1 + 2 w2 + w3 - 4 w)/3 style combination on whole registers:
// out = c0*a0 + c1*a1 + c2*a2 + c3*a3 (null operands skipped)
__global__ void k_axpy4(long long n, double* __restrict__ out, const double* a0, double c0,
                        const double* a1, double c1, const double* a2, double c2, const double* a3,
                        double c3) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = c0 * a0[i];
  if (a1) s += c1 * a1[i];
  if (a2) s += c2 * a2[i];
  if (a3) s += c3 * a3[i];
  out[i] = s;
}

// ---------------------------------------------------------------------
// launchers

template <int NDIM, int NDIR>
static void launch_nd(int which, int axis, const void* args, hipStream_t st, long long nthreads) {
  const int bs = 256;
  unsigned nb = (unsigned)((nthreads + bs - 1) / bs);
  if (which <= 1) nb = (nb + 7) / 8 * 8;  // multiple of 8 for xcd_block
  if (nb == 0) return;
  if (which == 0) {
    const AxisArgs& A = *(const AxisArgs*)args;
    if (axis == 0) hipLaunchKernelGGL((k_axis<NDIM, NDIR, 0>), dim3(nb), dim3(bs), 0, st, A);
    if constexpr (NDIM > 1)
      if (axis == 1) hipLaunchKernelGGL((k_axis<NDIM, NDIR, 1>), dim3(nb), dim3(bs), 0, st, A);
    if constexpr (NDIM > 2)
      if (axis == 2) hipLaunchKernelGGL((k_axis<NDIM, NDIR, 2>), dim3(nb), dim3(bs), 0, st, A);
  } else if (which == 1) {
    hipLaunchKernelGGL((k_final<NDIM, NDIR>), dim3(nb), dim3(bs), 0, st, *(const FinalArgs*)args);
  } else {
    hipLaunchKernelGGL((k_cmax<NDIM, NDIR>), dim3(nb), dim3(bs), 0, st, *(const CmaxArgs*)args);
  }
}

static void dispatch(int ndim, int ndir, int which, int axis, const void* args, hipStream_t st,
                     long long n) {
  if (ndim == 1 && ndir == 1) launch_nd<1, 1>(which, axis, args, st, n);
  else if (ndim == 1 && ndir == 2) launch_nd<1, 2>(which, axis, args, st, n);
  else if (ndim == 1 && ndir == 3) launch_nd<1, 3>(which, axis, args, st, n);
  else if (ndim == 2 && ndir == 2) launch_nd<2, 2>(which, axis, args, st, n);
  else if (ndim == 2 && ndir == 3) launch_nd<2, 3>(which, axis, args, st, n);
  else if (ndim == 3 && ndir == 3) launch_nd<3, 3>(which, axis, args, st, n);
}

static long long mesh_cells(const GridDesc& g) {
  long long n = 1;
  for (int d = 0; d < 3; d++) n *= (g.Mhi[d] - g.Mlo[d] + 1);
  return n;
}

void launch_axis(const AxisArgs& a, int axis, hipStream_t st) {
  dispatch(a.g.ndim, a.g.ndir, 0, axis, &a, st, mesh_cells(a.g));
}

void launch_final(const FinalArgs& a, hipStream_t st) {
  dispatch(a.g.ndim, a.g.ndir, 1, 0, &a, st, mesh_cells(a.g));
}

void launch_cmax(const CmaxArgs& a, hipStream_t st) {
  dispatch(a.g.ndim, a.g.ndir, 2, 0, &a, st, mesh_cells(a.g));
}

void launch_ghost(const GhostArgs& a, hipStream_t st) {
  const GridDesc& g = a.g;
  long long ncell = (long long)g.G[0] * g.G[1] * g.G[2];
  long long n = ncell - mesh_cells(g);
  if (n <= 0) return;
  const int bs = 256;
  hipLaunchKernelGGL(k_ghost, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st, a);
}

void launch_zero_u64(unsigned long long* p, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
}

void launch_axpy4(long long n, double* out, const double* a0, double c0, const double* a1,
                  double c1, const double* a2, double c2, const double* a3, double c3,
                  hipStream_t st) {
  const int bs = 256;
  hipLaunchKernelGGL(k_axpy4, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st, n, out, a0,
                     c0, a1, c1, a2, c2, a3, c3);
}

}  // namespace sac
