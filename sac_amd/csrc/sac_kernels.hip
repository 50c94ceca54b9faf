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
// kernel 1: setnu face ratios  r = md3R/md1R at faces j+1/2 along k
// (src/vacusr.viscosity.t:296-427); cells j in [Mlo-1, Mhi] along k and
// the mesh along the other axes.  q beyond the grid is the mirror value of
// the cell six inside (src/vacusr.viscosity.t:378-400).

template <int NDIM, int NDIR>
__global__ void __launch_bounds__(256) k_ratio(RatioArgs a) {
  const GridDesc& g = a.g;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = (d < NDIM) ? g.Mlo[d] - 1 : 0;
    ext[d] = (d < NDIM) ? g.Mhi[d] - g.Mlo[d] + 2 : 1;
  }
  if (t >= (long long)ext[0] * ext[1] * ext[2]) return;
  int c[3];
  decode_box(t, lo, ext, c);
  const long long fs = g.fs;
  const double* __restrict__ W = a.wct;
  const double* __restrict__ B = a.bg;
  constexpr int NP = 2 * NDIM + 1;
#pragma unroll
  for (int k = 0; k < NDIM; k++) {
    bool ok = true;
#pragma unroll
    for (int d = 0; d < NDIM; d++)
      if (d != k && c[d] < g.Mlo[d]) ok = false;
    if (!ok) continue;
    // q values at j-2..j+3 along k
    double q[NP][6];
#pragma unroll
    for (int s = 0; s < 6; s++) {
      int cc[3] = {c[0], c[1], c[2]};
      int jj = c[k] + s - 2;
      if (jj < 0) jj = g.Mlo[k] + 3;             // mirror below the grid
      if (jj > g.G[k] - 1) jj = g.Mhi[k] - 3;    // mirror above the grid
      cc[k] = jj;
      const long long i = lin(g, cc[0], cc[1], cc[2]);
      const double rho = W[i];
      const double rt = rho + B[fs + i];
      const double inv = 1.0 / rt;
      double m[NDIR], b[NDIR];
      double sb = 0.0, sm = 0.0;
#pragma unroll
      for (int cdir = 0; cdir < NDIR; cdir++) {
        m[cdir] = W[(1 + cdir) * fs + i];
        b[cdir] = W[(2 + NDIR + cdir) * fs + i];
        sb += b[cdir] * b[cdir];
        sm += m[cdir] * m[cdir];
      }
      const double e = W[(1 + NDIR) * fs + i];
      q[0][s] = rho;
      q[1][s] = e - 0.5 * (sb + sm * inv);
#pragma unroll
      for (int l = 0; l < NDIM; l++) q[2 + l][s] = m[l] * inv;
      int n = 2 + NDIM;
#pragma unroll
      for (int l = 0; l < NDIM; l++)
        if (l != k) q[n++][s] = b[l];
    }
    const long long io = lin(g, c[0], c[1], c[2]);
#pragma unroll
    for (int p = 0; p < NP; p++) {
      double d1[5], d3[3];
#pragma unroll
      for (int j = 0; j < 5; j++) d1[j] = fabs(q[p][j + 1] - q[p][j]);
#pragma unroll
      for (int j = 0; j < 3; j++)
        d3[j] = fabs(3.0 * (q[p][j + 2] - q[p][j + 1]) - (q[p][j + 3] - q[p][j]));
      double md3 = fmax(fmax(d3[0], d3[1]), d3[2]);
      double md1 = fmax(fmax(fmax(d1[0], d1[1]), fmax(d1[2], d1[3])), d1[4]);
      const double r = (md1 > 0.0) ? md3 / md1 : 0.0;
      a.ratio[(long long)(k * NP + p) * fs + io] = r;
    }
  }
}

// ---------------------------------------------------------------------
// kernel 2: fused stage update on the mesh.
//   R = -sum_d D4_d(F_d) + background energy terms + gravity + hyperdiffusion
//   out = base + a*R ; acc_out = acc_in + b*R (optional)
// plus max reductions of cmax_d(out) and cmax_d(out)/dx_d, and of every nu.

template <int NDIM, int NDIR>
__global__ void __launch_bounds__(256) k_stage(StageArgs a) {
  constexpr int NE = 2 + 2 * NDIR;
  constexpr int NP = 2 * NDIM + 1;
  const GridDesc& g = a.g;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = g.Mlo[d];
    ext[d] = g.Mhi[d] - g.Mlo[d] + 1;
  }
  const long long ncell = (long long)ext[0] * ext[1] * ext[2];
  const bool active = t < ncell;
  double redv[2 * NDIM + 1];
#pragma unroll
  for (int k = 0; k < 2 * NDIM + 1; k++) redv[k] = 0.0;

  if (active) {
    int c[3];
    decode_box(t, lo, ext, c);
    const long long fs = g.fs;
    const double* __restrict__ W = a.wct;
    const double* __restrict__ B = a.bg;
    const double gamma = a.gamma;
    const long long ic = lin(g, c[0], c[1], c[2]);
    double R[NE];
#pragma unroll
    for (int v = 0; v < NE; v++) R[v] = 0.0;

    Pt<NDIR> p0;
    load_pt<NDIR>(W, B, fs, ic, p0);
    const double rt0 = p0.rho + p0.rhob;
    const double inv0 = 1.0 / rt0;
    const double pTb0 = ptotal_bg<NDIR>(p0, gamma);
    double dxc[3];
#pragma unroll
    for (int d = 0; d < NDIM; d++) dxc[d] = a.dx[d][c[d]];

    // ---- flux divergence and background energy terms (centdiff4) ----
    // per axis k: points at c-2e_k, c-e_k, c+e_k, c+2e_k; the +-1
    // neighbours also supply the hyperdiffusion variables
    double nb_rho[NDIM][2], nb_qe[NDIM][2], nb_v[NDIM][NDIM][2], nb_b[NDIM][NDIM][2],
        nb_rt[NDIM][2], nb_m[NDIM][NDIM][2];
#pragma unroll
    for (int k = 0; k < NDIM; k++) {
      const bool edge4 = (c[k] <= g.Mlo[k] + 1) || (c[k] >= g.Mhi[k] - 1);
      double Fs[4][NE];
      double vs[4][NDIM];
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const int off = (s < 2) ? s - 2 : s - 1;  // -2,-1,+1,+2
        if (edge4 && (off == -2 || off == 2)) continue;
        const long long i = ic + (long long)off * g.st[k];
        Pt<NDIR> p;
        load_pt<NDIR>(W, B, fs, i, p);
        double F[NE];
        flux_pt<NDIR>(p, k, gamma, F);
#pragma unroll
        for (int v = 0; v < NE; v++) Fs[s][v] = F[v];
        const double rt = p.rho + p.rhob;
        const double inv = 1.0 / rt;
#pragma unroll
        for (int l = 0; l < NDIM; l++) vs[s][l] = p.m[l] * inv;
        if (off == -1 || off == 1) {
          const int sd = (off == -1) ? 0 : 1;
          double sb = 0.0, sm = 0.0;
#pragma unroll
          for (int cdir = 0; cdir < NDIR; cdir++) {
            sb += p.b[cdir] * p.b[cdir];
            sm += p.m[cdir] * p.m[cdir];
          }
          nb_rho[k][sd] = p.rho;
          nb_qe[k][sd] = p.e - 0.5 * (sb + sm * inv);
          nb_rt[k][sd] = rt;
#pragma unroll
          for (int l = 0; l < NDIM; l++) {
            nb_v[k][l][sd] = p.m[l] * inv;
            nb_m[k][l][sd] = p.m[l];
            nb_b[k][l][sd] = p.b[l];
          }
        }
      }
      if (!edge4) {
        const double inv12 = 1.0 / dxc[k] / 12.0;
#pragma unroll
        for (int v = 0; v < NE; v++) {
          const double D = -(Fs[3][v] - 8.0 * (Fs[2][v] - Fs[1][v]) - Fs[0][v]) * inv12;
          R[v] -= D;
        }
        // background energy terms with derivative along k
        double Dv[NDIM];
#pragma unroll
        for (int l = 0; l < NDIM; l++)
          Dv[l] = -(vs[3][l] - 8.0 * (vs[2][l] - vs[1][l]) - vs[0][l]) * inv12;
        double re = -Dv[k] * pTb0;
#pragma unroll
        for (int l = 0; l < NDIM; l++) re += p0.bg[k] * p0.bg[l] * Dv[l];
        R[1 + NDIR] += re;
      }
    }

    // ---- gravity on the perturbed density (addsource_grav) ----
    {
      double rg = 0.0;
#pragma unroll
      for (int d = 0; d < NDIM; d++) {
        if (a.grav_on[d]) {
          R[1 + d] += a.grav[d] * p0.rho;
          rg += a.grav[d] * p0.rho * p0.m[d] * inv0;
        }
      }
      R[1 + NDIR] += rg;
    }

    // ---- hyperdiffusion (addsource_visc) ----
    double numax = 0.0;
    if (a.visc_on) {
      double qe0;
      {
        double sb = 0.0, sm = 0.0;
#pragma unroll
        for (int cdir = 0; cdir < NDIR; cdir++) {
          sb += p0.b[cdir] * p0.b[cdir];
          sm += p0.m[cdir] * p0.m[cdir];
        }
        qe0 = p0.e - 0.5 * (sb + sm * inv0);
      }
      double ctot[NDIM];
#pragma unroll
      for (int k = 0; k < NDIM; k++) ctot[k] = red_get(a.ctot_src, k);
      const double* __restrict__ RT = a.ratio;

#pragma unroll
      for (int k = 0; k < NDIM; k++) {
        // gradient1L/1R are zero at the first/last mesh cell of a
        // non-periodic side
        const bool z1 = (c[k] == g.Mlo[k] && !g.per_lo[k]) || (c[k] == g.Mhi[k] && !g.per_hi[k]);
        const long long im = ic - g.st[k];
        const double dxk = dxc[k];
        double nuR[NP], nuL[NP];
#pragma unroll
        for (int p = 0; p < NP; p++) {
          const double ch = (p >= 2 && p < 2 + NDIM) ? 0.4 : 0.04;
          const double K = ctot[k] * ch;
          nuR[p] = K * RT[(long long)(k * NP + p) * fs + ic] * dxk;
          nuL[p] = K * RT[(long long)(k * NP + p) * fs + im] * dxk;
          numax = fmax(numax, fmax(nuR[p], nuL[p]));
        }
        if (!z1) {
          // rho
          {
            const double gL = (p0.rho - nb_rho[k][0]) / dxk;
            const double gR = (nb_rho[k][1] - p0.rho) / dxk;
            R[0] += (nuR[0] * gR - nuL[0] * gL) / dxk;
          }
          // thermal energy code 173
          {
            const double gL = (qe0 - nb_qe[k][0]) / dxk;
            const double gR = (nb_qe[k][1] - qe0) / dxk;
            R[1 + NDIR] += (nuR[1] * gR - nuL[1] * gL) / dxk;
          }
          // momentum, i==k branch (twice for l==k)
          const double rhoL = (rt0 + nb_rt[k][0]) / 2.0;
          const double rhoR = (nb_rt[k][1] + rt0) / 2.0;
#pragma unroll
          for (int l = 0; l < NDIM; l++) {
            const double v0 = p0.m[l] * inv0;
            const double tL = nuL[2 + l] * ((v0 - nb_v[k][l][0]) / dxk);
            const double tR = nuR[2 + l] * ((nb_v[k][l][1] - v0) / dxk);
            const double mL = (p0.m[l] + nb_m[k][l][0]) / 2.0;
            const double mR = (nb_m[k][l][1] + p0.m[l]) / 2.0;
            const double dm = (rhoR * tR - rhoL * tL) / dxk / 2.0;
            const double de = (mR * tR - mL * tL) / dxk / 2.0;
            const double mult = (l == k) ? 2.0 : 1.0;
            R[1 + l] += mult * dm;
            R[1 + NDIR] += mult * de;
          }
          // magnetic field, m==k branch, l != k
          int pb = 2 + NDIM;
#pragma unroll
          for (int l = 0; l < NDIM; l++) {
            if (l == k) continue;
            const double b0 = p0.b[l];
            const double tL = nuL[pb] * ((b0 - nb_b[k][l][0]) / dxk);
            const double tR = nuR[pb] * ((nb_b[k][l][1] - b0) / dxk);
            const double BL = (b0 + nb_b[k][l][0]) / 2.0;
            const double BR = (nb_b[k][l][1] + b0) / 2.0;
            R[2 + NDIR + l] += (tR - tL) / dxk;
            R[1 + NDIR] += (tR * BR - tL * BL) / dxk;
            pb++;
          }
        }
      }

      // cross terms: diffusion along k, derivative along l (l != k)
#pragma unroll
      for (int k = 0; k < NDIM; k++) {
#pragma unroll
        for (int l = 0; l < NDIM; l++) {
          if (l == k) continue;
          const bool zk = (c[k] == g.Mlo[k] && !g.per_lo[k]) || (c[k] == g.Mhi[k] && !g.per_hi[k]);
          const bool zl = (c[l] == g.Mlo[l] && !g.per_lo[l]) || (c[l] == g.Mhi[l] && !g.per_hi[l]);
          if (zl) continue;  // gradient1 along l is zero at c
          // pair indices: momentum m_l along k, field b_l along k
          const int pm = 2 + l;
          const int pbf = 2 + NDIM + (l < k ? l : l - 1);
          double Tm[2], Tb[2], rts[2], mks[2], bks[2];
#pragma unroll
          for (int sd = 0; sd < 2; sd++) {
            const int sl = sd ? 1 : -1;
            const int cl = c[l] + sl;
            rts[sd] = nb_rt[l][sd];
            mks[sd] = nb_m[l][k][sd];
            bks[sd] = nb_b[l][k][sd];
            Tm[sd] = 0.0;
            Tb[sd] = 0.0;
            const bool inmesh = (cl >= g.Mlo[l]) && (cl <= g.Mhi[l]);
            if (!inmesh || zk) continue;
            const long long ip = ic + (long long)sl * g.st[l];
            const long long ipp = ip + g.st[k];
            const long long ipm = ip - g.st[k];
            // v_l and b_l at the two diagonal neighbours
            const double rtp = W[ipp] + B[fs + ipp];
            const double rtm = W[ipm] + B[fs + ipm];
            const double vlp = W[(1 + l) * fs + ipp] / rtp;
            const double vlm = W[(1 + l) * fs + ipm] / rtm;
            const double blp = W[(2 + NDIR + l) * fs + ipp];
            const double blm = W[(2 + NDIR + l) * fs + ipm];
            const double dxk = dxc[k];
            const double Kv = ctot[k] * 0.4;
            const double Kb = ctot[k] * 0.04;
            const double nuLm = Kv * RT[(long long)(k * NP + pm) * fs + ip - g.st[k]] * dxk;
            const double nuRm = Kv * RT[(long long)(k * NP + pm) * fs + ip] * dxk;
            const double nuLb = Kb * RT[(long long)(k * NP + pbf) * fs + ip - g.st[k]] * dxk;
            const double nuRb = Kb * RT[(long long)(k * NP + pbf) * fs + ip] * dxk;
            Tm[sd] = (vlp - vlm) / dxk / 2.0 * (nuLm + nuRm) / 2.0 / 2.0;
            Tb[sd] = (blp - blm) / dxk / 2.0 * (nuLb + nuRb) / 2.0;
          }
          const double dxl = dxc[l];
          // m_k += d_l(rho_t T), e += d_l(m_k T)
          R[1 + k] += (rts[1] * Tm[1] - rts[0] * Tm[0]) / dxl / 2.0;
          R[1 + NDIR] += (mks[1] * Tm[1] - mks[0] * Tm[0]) / dxl / 2.0;
          // b_k -= d_l(Tb), e -= d_l(Tb b_k)
          R[2 + NDIR + k] -= (Tb[1] - Tb[0]) / dxl / 2.0;
          R[1 + NDIR] -= (Tb[1] * bks[1] - Tb[0] * bks[0]) / dxl / 2.0;
        }
      }
    }
    redv[2 * NDIM] = numax;

    // ---- register update ----
    Pt<NDIR> po;
    {
      const double* __restrict__ base = a.base;
      double* __restrict__ out = a.out;
      double vals[NE];
#pragma unroll
      for (int v = 0; v < NE; v++) {
        vals[v] = base[v * fs + ic] + a.qa * R[v];
        out[v * fs + ic] = vals[v];
      }
      if (a.acc_out) {
#pragma unroll
        for (int v = 0; v < NE; v++) a.acc_out[v * fs + ic] = a.acc_in[v * fs + ic] + a.qb * R[v];
      }
      po.rho = vals[0];
#pragma unroll
      for (int cdir = 0; cdir < NDIR; cdir++) {
        po.m[cdir] = vals[1 + cdir];
        po.b[cdir] = vals[2 + NDIR + cdir];
        po.bg[cdir] = p0.bg[cdir];
      }
      po.e = vals[1 + NDIR];
      po.eb = p0.eb;
      po.rhob = p0.rhob;
    }
    if (a.reduce_out) {
#pragma unroll
      for (int d = 0; d < NDIM; d++) {
        const double cm = cmax_pt<NDIR>(po, d, gamma);
        redv[d] = cm;
        redv[NDIM + d] = cm / dxc[d];
      }
    }
  }
  // reductions (all threads take part)
  {
    double r1[2 * NDIM];
#pragma unroll
    for (int k = 0; k < 2 * NDIM; k++) r1[k] = redv[k];
    if (a.reduce_out) block_max_atomic<2 * NDIM>(r1, a.red_out);
    double r2[1] = {redv[2 * NDIM]};
    if (a.visc_on) block_max_atomic<1>(r2, a.red_visc);
  }
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
static void launch_all(int which, const void* args, hipStream_t st, long long nthreads) {
  const int bs = 256;
  const unsigned nb = (unsigned)((nthreads + bs - 1) / bs);
  if (nb == 0) return;
  switch (which) {
    case 0:
      hipLaunchKernelGGL((k_ratio<NDIM, NDIR>), dim3(nb), dim3(bs), 0, st,
                         *(const RatioArgs*)args);
      break;
    case 1:
      hipLaunchKernelGGL((k_stage<NDIM, NDIR>), dim3(nb), dim3(bs), 0, st,
                         *(const StageArgs*)args);
      break;
    case 2:
      hipLaunchKernelGGL((k_cmax<NDIM, NDIR>), dim3(nb), dim3(bs), 0, st,
                         *(const CmaxArgs*)args);
      break;
  }
}

static void dispatch(int ndim, int ndir, int which, const void* args, hipStream_t st,
                     long long n) {
  if (ndim == 1 && ndir == 1) launch_all<1, 1>(which, args, st, n);
  else if (ndim == 1 && ndir == 2) launch_all<1, 2>(which, args, st, n);
  else if (ndim == 1 && ndir == 3) launch_all<1, 3>(which, args, st, n);
  else if (ndim == 2 && ndir == 2) launch_all<2, 2>(which, args, st, n);
  else if (ndim == 2 && ndir == 3) launch_all<2, 3>(which, args, st, n);
  else if (ndim == 3 && ndir == 3) launch_all<3, 3>(which, args, st, n);
}

static long long mesh_cells(const GridDesc& g) {
  long long n = 1;
  for (int d = 0; d < 3; d++) n *= (g.Mhi[d] - g.Mlo[d] + 1);
  return n;
}

void launch_ratio(const RatioArgs& a, hipStream_t st) {
  long long n = 1;
  for (int d = 0; d < 3; d++)
    n *= (d < a.g.ndim) ? (a.g.Mhi[d] - a.g.Mlo[d] + 2) : 1;
  dispatch(a.g.ndim, a.g.ndir, 0, &a, st, n);
}

void launch_stage(const StageArgs& a, hipStream_t st) {
  dispatch(a.g.ndim, a.g.ndir, 1, &a, st, mesh_cells(a.g));
}

void launch_cmax(const CmaxArgs& a, hipStream_t st) {
  dispatch(a.g.ndim, a.g.ndir, 2, &a, st, mesh_cells(a.g));
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
