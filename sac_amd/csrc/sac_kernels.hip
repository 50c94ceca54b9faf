// sac_kernels.hip -- HIP kernels (gfx950) of the SAC explicit
// perturbation-MHD stage: per-axis line work (fluxes, setnu, hyperdiffusion),
// the combining update (cross terms, gravity, Runge-Kutta register update and
// the cmax reductions of the result), ghost-cell fill and reductions.
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
#include "sac_device.h"
#include "sac_line.h"
#include "sac_plane.h"

// SAC_LINE_TILES=0 keeps the line-group kernel k_axis for every grid (the
// default runs 2-D and 3-D grids through the LDS line tiles, k_line)
#ifndef SAC_LINE_TILES
#define SAC_LINE_TILES 1
#endif

namespace sac {

// kernel 1: line groups of axis K (see axis_group)
template <int NDIM, int NDIR, int K, typename S, typename C, bool LEARNED = false>
__global__ void __launch_bounds__(256) k_axis(AxisArgs a) {
  constexpr int CK = LineGroup<K>::CK;
  const GridDesc& g = a.g;
  const long long t = xcd_block(a.xcd) * blockDim.x + threadIdx.x;
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = g.Mlo[d];
    ext[d] = g.Mhi[d] - g.Mlo[d] + 1;
  }
  // groups of CK cells along K
  ext[K] = (ext[K] + CK - 1) / CK;
  const bool active = t < (long long)ext[0] * ext[1] * ext[2];
  C numax = 0;
  if (active) {
    int c[3];
    {
      long long u = t;
      c[0] = (int)(u % ext[0]);
      u /= ext[0];
      c[1] = (int)(u % ext[1]);
      u /= ext[1];
      c[2] = (int)u;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) c[d] = lo[d] + ((d == K) ? c[d] * CK : c[d]);
    axis_group<NDIM, NDIR, K, CK, S, C, LEARNED>(a, g, c, numax);
  }
  if (a.visc_on) {
    double r2[1] = {(double)numax};
    block_max_atomic<1>(r2, a.red_visc);
  }
}

// kernel 1, 2-D and 3-D grids: LDS-staged line tiles of axis K (see line_tile)
template <int NDIM, int NDIR, int K, typename S, typename C, bool LEARNED = false,
          bool DIVB = false>
__global__ void __launch_bounds__((line::Tile<K, C>::T)) k_line(AxisArgs a, int seg) {
  __shared__ C L[line::lds_elems<NDIM, NDIR, K, C, DIVB>()];
  C numax = 0;
  line::line_tile<NDIM, NDIR, K, S, C, LEARNED, DIVB>(a, seg, L, numax);
  if (a.visc_on) {
    double r2[1] = {(double)numax};
    block_max_atomic<1>(r2, a.red_visc);
  }
}

// kernel 1, 3-D grids: the x and y work of z-plane tiles (see plane_tile)
template <typename S, typename C>
__global__ void __launch_bounds__(plane::NT) k_plane(AxisArgs a) {
  __shared__ plane::Smem<C> sm;
  C numax = 0;
  plane::plane_tile<S, C>(a, sm, numax);
  if (a.visc_on) {
    double r2[1] = {(double)numax};
    block_max_atomic<1>(r2, a.red_visc);
  }
}

// kernel 1, 1-D grids in FP64: LDS line tiles of 256 cells (line1_tile)
template <int NDIR>
__global__ void __launch_bounds__(256) k_line1(AxisArgs a) {
  __shared__ double L[line1::LV<NDIR> * (256 + 6)];
  double numax = 0.0;
  line1_tile<256, NDIR>(a.g, (const double*)a.wct, a.bg, a.dx[0], (double*)a.R, a.gamma, a.visc_on,
                  a.visc_on ? red_get(a.ctot_src, 0) : 0.0, a.g.Mlo[0] + 256 * (int)blockIdx.x, L,
                  numax);
  if (a.visc_on) {
    double r2[1] = {numax};
    block_max_atomic<1>(r2, a.red_visc);
  }
}

// segment length along K: whole lines unless that leaves fewer than 2048
// blocks (then halves, down to 32 cells)
template <int NDIM, int K, typename C>
static int line_segment(const GridDesc& g) {
  const int ext = g.Mhi[K] - g.Mlo[K] + 1;
  const long long lines = line::nblocks<NDIM, K, C>(g, ext);
  int nseg = 1;
  while (lines * nseg < 2048 && ext / (2 * nseg) >= 32) nseg *= 2;
  return (ext + nseg - 1) / nseg;
}

template <int NDIM, int NDIR, int K, typename S, typename C, bool LEARNED = false>
static void launch_line(const AxisArgs& A, hipStream_t st) {
  const int seg = line_segment<NDIM, K, C>(A.g);
  long long nb = line::nblocks<NDIM, K, C>(A.g, seg);
  nb = (nb + 15) / 16 * 16;  // tile pairing over the XCDs (see line_tile)
  if (nb <= 0) return;
  // Powell's sources need one more ring value per point: a kernel variant
  if (A.divbfix)
    hipLaunchKernelGGL((k_line<NDIM, NDIR, K, S, C, LEARNED, true>), dim3((unsigned)nb),
                       dim3(line::Tile<K, C>::T), 0, st, A, seg);
  else
    hipLaunchKernelGGL((k_line<NDIM, NDIR, K, S, C, LEARNED, false>), dim3((unsigned)nb),
                       dim3(line::Tile<K, C>::T), 0, st, A, seg);
}

// kernel 2: the combining update of every mesh cell (see final_cells), FV
// consecutive cells along x per thread; RK4L: RK4's last stage
template <int NDIM, int NDIR, typename S, typename C>
struct FinalV {
  // FP32 storage: two cells per thread, so that neighbouring cells' loads
  // pair into 8-byte ones (FP64 values already are)
  static constexpr int V = (sizeof(S) == 4) ? 2 : 1;
};

template <int NDIM, int NDIR, typename S, typename C, bool RK4L>
__global__ void __launch_bounds__(256) k_final(FinalArgs a) {
  constexpr int V = FinalV<NDIM, NDIR, S, C>::V;
  const GridDesc& g = a.g;
  const long long t = xcd_block(a.xcd) * blockDim.x + threadIdx.x;
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = g.Mlo[d];
    ext[d] = g.Mhi[d] - g.Mlo[d] + 1;
  }
  ext[0] = (ext[0] + V - 1) / V;
  const bool active = t < (long long)ext[0] * ext[1] * ext[2];
  double redv[2 * NDIM];
#pragma unroll
  for (int k = 0; k < 2 * NDIM; k++) redv[k] = 0.0;
  if (active) {
    int c[3];
    decode_box(t, lo, ext, c);
    c[0] = lo[0] + V * (c[0] - lo[0]);
    const int nv = (V == 1) ? 1 : min(V, g.Mhi[0] - c[0] + 1);
    final_cells<NDIM, NDIR, S, C, V, RK4L>(a, g, c, nv, redv);
  }
  block_max_atomic<2 * NDIM>(redv, a.red_out);
}

// ---------------------------------------------------------------------
// kernel 3: cmax reductions of a register (getdt_courant / setnu c_tot)

template <int NDIM, int NDIR, typename S>
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
    Pt<NDIR, double> p;
    load_pt<NDIR, S, double>((const S*)a.w, a.bg, g.fs, lin(g, c[0], c[1], c[2]), p);
    double cm[NDIM];
    cmax_all<NDIM, NDIR>(p, a.gamma, cm);
#pragma unroll
    for (int d = 0; d < NDIM; d++) {
      r[d] = cm[d];
      r[NDIM + d] = cm[d] / a.dx[d][c[d]];
    }
  }
  block_max_atomic<2 * NDIM>(r, a.red_out);
}

// kernel 4: ghost-cell fill (see ghost_value)
// cell t of the ghost shell = part A (ghost along axis 2 [0-based], any
// i0,i1) + part B (mesh along 2, ghost along 1, any i0) + part C (mesh along
// 2 and 1, ghost along 0); false past the shell
__device__ __forceinline__ bool shell_cell(const GridDesc& g, long long t, int (&c)[3]) {
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
    return false;
  }
  return true;
}

template <typename S, bool GENERAL>
__global__ void __launch_bounds__(256) k_ghost(GhostArgs a) {
  const GridDesc& g = a.g;
  int c[3];
  if (!shell_cell(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, c)) return;
  const long long io = lin(g, c[0], c[1], c[2]);
  S* f = (S*)a.f;
  for (int v = 0; v < a.nvar; v++) {
    const S val = ghost_value<S, GENERAL>(a, v, c[0], c[1], c[2]);
    f[(long long)v * g.fs + io] = val;
  }
}

// stored values of the fixed / fixed1 / grad1 ghosts of the filled field a.f
template <typename S>
__global__ void __launch_bounds__(256) k_ghost_store(GhostArgs a, S* __restrict__ store) {
  const GridDesc& g = a.g;
  int c[3];
  if (!shell_cell(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, c)) return;
  const long long io = lin(g, c[0], c[1], c[2]);
  for (int v = 0; v < a.nvar; v++)
    store[(long long)v * g.fs + io] = ghost_stored<S>(a, v, c[0], c[1], c[2]);
}

// ---------------------------------------------------------------------
// kernel 4b: grids with 'zero' boundaries fill their ghost regions one
// after another, as the reference does (src/vacgrid.t:195-512), because
// every 'zero' fill converts the whole grid to primitive variables and back
// (src/vacgrid.t:294-345), which moves every cell by rounding: no ghost value
// can be traced back to mesh cells as k_ghost does.

template <typename S>
__global__ void __launch_bounds__(256) k_region_fill(RegionArgs a) {
  const GridDesc& g = a.g;
  const int D = a.D;
  const int nlay = (a.side ? g.G[D] - 1 - g.Mhi[D] : g.Mlo[D]);  // ghost layers
  int ext[3] = {g.G[0], g.G[1], g.G[2]};
  ext[D] = nlay;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ext[0] * ext[1] * ext[2]) return;
  const int lo3[3] = {0, 0, 0};
  int c[3];
  decode_box(t, lo3, ext, c);
  const int lo = a.side ? g.Mhi[D] + 1 : 0, hi = lo + nlay - 1;  // the region along D
  c[D] += lo;
  S* f = (S*)a.f;
  auto at = [&](int cd) {
    int e[3] = {c[0], c[1], c[2]};
    e[D] = cd;
    return f[lin(g, e[0], e[1], e[2])];
  };
  const int N = g.Mhi[D] - g.Mlo[D] + 1;
  const int ixe = a.side ? g.Mhi[D] : g.Mlo[D];  // edge cell
  const long long i = lin(g, c[0], c[1], c[2]);
  switch (a.code) {
    case BC_PERIODIC:
      f[i] = at(c[D] + (a.side ? -N : N));
      break;
    case BC_SYMM:
    case BC_SYMM0:
      f[i] = at(a.side ? (2 * g.Mhi[D] + 1 - c[D]) : (2 * g.Mlo[D] - 1 - c[D]));
      break;
    case BC_ASYMM:
      f[i] = -at(a.side ? (2 * g.Mhi[D] + 1 - c[D]) : (2 * g.Mlo[D] - 1 - c[D]));
      break;
    case BC_CONTCD4:
    case BC_ZERO: {  // upper ghosts lo <- lo-2, hi <- lo-3; lower lo <- hi+3, hi <- hi+2
      const int src = (c[D] == lo) ? (a.side ? lo - 2 : hi + 3) : (a.side ? lo - 3 : hi + 2);
      if (c[D] == lo || c[D] == hi) f[i] = at(src);
      break;
    }
    case BC_FIXED:
      if (!a.keep_fixed) f[i] = at(ixe);  // the first fill is cont's
      break;
    case BC_FIXED1:
    case BC_GRAD1:
      if (a.keep_fixed) break;
      [[fallthrough]];
    case BC_CONT1: {
#pragma clang fp contract(off)
      const int k = abs(c[D] - ixe);
      const int ixf = a.side ? ixe - 1 : ixe + 1;
      f[i] = (S)((double)(k + 1) * (double)at(ixe) - (double)k * (double)at(ixf));
      break;
    }
    default:  // cont
      f[i] = at(ixe);
  }
}

// primitive: e <- p_thermal (getpthermal, src/vacphys0.t.mhd:163-183),
// m <- m/rho_t; conserve: e <- p/(gamma-1) + half*(rho_t*sum v^2 + sum b^2)
// + sum b*bg, m <- rho_t*v.  Every cell of the grid, ghosts included.
template <int NDIR, typename S>
__global__ void __launch_bounds__(256) k_prim_cons(GridDesc g, S* w, const double* bg,
                                                   double gamma, int to_primitive) {
#pragma clang fp contract(off)
  const long long ncell = (long long)g.G[0] * g.G[1] * g.G[2];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  int c[3];
  const int lo3[3] = {0, 0, 0};
  decode_box(t, lo3, g.G, c);
  const long long i = lin(g, c[0], c[1], c[2]), fs = g.fs;
  const double rt = (double)w[i] + bg[fs + i];
  double m[NDIR], sb = 0.0, sbb = 0.0;
#pragma unroll
  for (int cd = 0; cd < NDIR; cd++) {
    m[cd] = (double)w[(1 + cd) * fs + i];
    const double b = (double)w[(2 + NDIR + cd) * fs + i];
    sb = (cd == 0) ? b * b : sb + b * b;
    sbb = (cd == 0) ? b * bg[(2 + cd) * fs + i] : sbb + b * bg[(2 + cd) * fs + i];
  }
  const double e = (double)w[(1 + NDIR) * fs + i];
  if (to_primitive) {
    double sm = 0.0;
#pragma unroll
    for (int cd = 0; cd < NDIR; cd++) sm = (cd == 0) ? m[cd] * m[cd] : sm + m[cd] * m[cd];
    double p = 0.5 * sm / rt;
    p = p + 0.5 * sb + sbb;
    w[(1 + NDIR) * fs + i] = (S)((gamma - 1.0) * (e - p));
#pragma unroll
    for (int cd = 0; cd < NDIR; cd++) w[(1 + cd) * fs + i] = (S)(m[cd] / rt);
  } else {
    double sv = 0.0;  // the momentum slots hold v
#pragma unroll
    for (int cd = 0; cd < NDIR; cd++) sv = (cd == 0) ? m[cd] * m[cd] : sv + m[cd] * m[cd];
    w[(1 + NDIR) * fs + i] = (S)(e / (gamma - 1.0) + 0.5 * (rt * sv + sb) + sbb);
#pragma unroll
    for (int cd = 0; cd < NDIR; cd++) w[(1 + cd) * fs + i] = (S)(rt * m[cd]);
  }
}

// ---------------------------------------------------------------------
// small helpers

__global__ void k_zero_u64(unsigned long long* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// out = c0*a0 + c1*a1 + c2*a2 + c3*a3 in FP64 (null operands skipped)
template <typename S>
__global__ void k_axpy4(long long n, S* __restrict__ out, const S* a0, double c0, const S* a1,
                        double c1, const S* a2, double c2, const S* a3, double c3) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = c0 * (double)a0[i];
  if (a1) s += c1 * (double)a1[i];
  if (a2) s += c2 * (double)a2[i];
  if (a3) s += c3 * (double)a3[i];
  out[i] = (S)s;
}

// type conversion of n elements (FP64 staging <-> FP32 registers)
template <typename D, typename S>
__global__ void k_convert(long long n, D* __restrict__ out, const S* __restrict__ in) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (D)in[i];
}

// ---------------------------------------------------------------------
// launchers

template <int NDIM, int NDIR, typename S, typename C>
static void launch_nd(int which, int axis, const void* args, hipStream_t st, long long nthreads) {
  const int bs = 256;
  unsigned nb = (unsigned)((nthreads + bs - 1) / bs);
  if (which <= 1) nb = (nb + 7) / 8 * 8;  // multiple of 8 for xcd_block
  if (nb == 0) return;
  if (which == 0) {
    const AxisArgs& A = *(const AxisArgs*)args;
    if (A.RF) {  // learned closure (FP64 storage only)
      if constexpr (sizeof(S) == sizeof(double)) {
        if constexpr (NDIM > 1 && SAC_LINE_TILES) {
          if (axis == 0) launch_line<NDIM, NDIR, 0, S, C, true>(A, st);
          if (axis == 1) launch_line<NDIM, NDIR, 1, S, C, true>(A, st);
          if constexpr (NDIM > 2)
            if (axis == 2) launch_line<NDIM, NDIR, 2, S, C, true>(A, st);
        } else {
          if (axis == 0)
            hipLaunchKernelGGL((k_axis<NDIM, NDIR, 0, S, C, true>), dim3(nb), dim3(bs), 0, st, A);
          if constexpr (NDIM > 1)
            if (axis == 1)
              hipLaunchKernelGGL((k_axis<NDIM, NDIR, 1, S, C, true>), dim3(nb), dim3(bs), 0, st, A);
          if constexpr (NDIM > 2)
            if (axis == 2)
              hipLaunchKernelGGL((k_axis<NDIM, NDIR, 2, S, C, true>), dim3(nb), dim3(bs), 0, st, A);
        }
      }
      return;
    }
    if constexpr (NDIM == 1 && sizeof(S) == sizeof(double) && SAC_LINE_TILES) {
      const int n = A.g.Mhi[0] - A.g.Mlo[0] + 1;
      hipLaunchKernelGGL(k_line1<NDIR>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, A);
      return;
    }
    if constexpr (NDIM > 1 && SAC_LINE_TILES) {
      if (axis == 0) launch_line<NDIM, NDIR, 0, S, C>(A, st);
      if (axis == 1) launch_line<NDIM, NDIR, 1, S, C>(A, st);
      if constexpr (NDIM > 2)
        if (axis == 2) launch_line<NDIM, NDIR, 2, S, C>(A, st);
      return;
    }
    if (axis == 0) hipLaunchKernelGGL((k_axis<NDIM, NDIR, 0, S, C>), dim3(nb), dim3(bs), 0, st, A);
    if constexpr (NDIM > 1)
      if (axis == 1)
        hipLaunchKernelGGL((k_axis<NDIM, NDIR, 1, S, C>), dim3(nb), dim3(bs), 0, st, A);
    if constexpr (NDIM > 2)
      if (axis == 2)
        hipLaunchKernelGGL((k_axis<NDIM, NDIR, 2, S, C>), dim3(nb), dim3(bs), 0, st, A);
  } else if (which == 1) {
    const FinalArgs& f = *(const FinalArgs*)args;
    if (f.w1)
      hipLaunchKernelGGL((k_final<NDIM, NDIR, S, C, true>), dim3(nb), dim3(bs), 0, st, f);
    else
      hipLaunchKernelGGL((k_final<NDIM, NDIR, S, C, false>), dim3(nb), dim3(bs), 0, st, f);
  } else {
    hipLaunchKernelGGL((k_cmax<NDIM, NDIR, S>), dim3(nb), dim3(bs), 0, st,
                       *(const CmaxArgs*)args);
  }
}

template <typename S, typename C>
static void dispatch_t(int ndim, int ndir, int which, int axis, const void* args, hipStream_t st,
                       long long n) {
  if (ndim == 1 && ndir == 1) launch_nd<1, 1, S, C>(which, axis, args, st, n);
  else if (ndim == 1 && ndir == 2) launch_nd<1, 2, S, C>(which, axis, args, st, n);
  else if (ndim == 1 && ndir == 3) launch_nd<1, 3, S, C>(which, axis, args, st, n);
  else if (ndim == 2 && ndir == 2) launch_nd<2, 2, S, C>(which, axis, args, st, n);
  else if (ndim == 2 && ndir == 3) launch_nd<2, 3, S, C>(which, axis, args, st, n);
  else if (ndim == 3 && ndir == 3) launch_nd<3, 3, S, C>(which, axis, args, st, n);
}

static void dispatch(int prec, int ndim, int ndir, int which, int axis, const void* args,
                     hipStream_t st, long long n) {
  if (prec == PREC_MIXED)
    dispatch_t<float, float>(ndim, ndir, which, axis, args, st, n);
  else
    dispatch_t<double, double>(ndim, ndir, which, axis, args, st, n);
}

static long long mesh_cells(const GridDesc& g) {
  long long n = 1;
  for (int d = 0; d < 3; d++) n *= (g.Mhi[d] - g.Mlo[d] + 1);
  return n;
}

int line_tiles_built() { return SAC_LINE_TILES ? 1 : 0; }

int launch_axis(const AxisArgs& a, int axis, hipStream_t st) {
  const GridDesc& g = a.g;
  long long n = 1;
  for (int d = 0; d < 3; d++) {
    long long e = g.Mhi[d] - g.Mlo[d] + 1;
    if (d == axis) {
      const int ck = (axis == 0) ? LineGroup<0>::CK : LineGroup<1>::CK;
      e = (e + ck - 1) / ck;
    }
    n *= e;
  }
  dispatch(a.prec, g.ndim, g.ndir, 0, axis, &a, st, n);
  const bool line = SAC_LINE_TILES && g.ndim > 1;
  return (line && a.cross_last && a.visc_on && axis == g.ndim - 1) ? 1 : 0;
}

void launch_plane(const AxisArgs& a, hipStream_t st) {
  long long nb = plane::nblocks(a.g);
  const int xcd = a.xcd > 1 ? a.xcd : 1;
  nb = (nb + xcd - 1) / xcd * xcd;  // xcd_block needs a multiple of the XCD count
  if (nb <= 0) return;
  if (a.prec == PREC_MIXED)
    hipLaunchKernelGGL((k_plane<float, float>), dim3((unsigned)nb), dim3(plane::NT), 0, st, a);
  else
    hipLaunchKernelGGL((k_plane<double, double>), dim3((unsigned)nb), dim3(plane::NT), 0, st, a);
}

void launch_final(const FinalArgs& a, hipStream_t st) {
  // threads: mesh cells, two per thread along x in FP32 storage
  const GridDesc& g = a.g;
  const int V = (a.prec == PREC_MIXED) ? 2 : 1;
  const long long n = (long long)((g.Mhi[0] - g.Mlo[0] + 1 + V - 1) / V) *
                      (g.Mhi[1] - g.Mlo[1] + 1) * (g.Mhi[2] - g.Mlo[2] + 1);
  dispatch(a.prec, g.ndim, g.ndir, 1, 0, &a, st, n);
}

void launch_cmax(const CmaxArgs& a, hipStream_t st) {
  dispatch(a.prec, a.g.ndim, a.g.ndir, 2, 0, &a, st, mesh_cells(a.g));
}

void launch_ghost(const GhostArgs& a, hipStream_t st) {
  const GridDesc& g = a.g;
  long long ncell = (long long)g.G[0] * g.G[1] * g.G[2];
  long long n = ncell - mesh_cells(g);
  if (n <= 0) return;
  const int bs = 256;
  const dim3 grid((unsigned)((n + bs - 1) / bs));
  if (a.prec == PREC_MIXED) {
    if (a.general)
      hipLaunchKernelGGL((k_ghost<float, true>), grid, dim3(bs), 0, st, a);
    else
      hipLaunchKernelGGL((k_ghost<float, false>), grid, dim3(bs), 0, st, a);
  } else {
    if (a.general)
      hipLaunchKernelGGL((k_ghost<double, true>), grid, dim3(bs), 0, st, a);
    else
      hipLaunchKernelGGL((k_ghost<double, false>), grid, dim3(bs), 0, st, a);
  }
}

void launch_region_fill(const RegionArgs& a, hipStream_t st) {
  const GridDesc& g = a.g;
  const int nlay = a.side ? g.G[a.D] - 1 - g.Mhi[a.D] : g.Mlo[a.D];
  long long n = (long long)g.G[0] * g.G[1] * g.G[2] / g.G[a.D] * nlay;
  if (n <= 0) return;
  const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
  if (a.prec == PREC_MIXED)
    hipLaunchKernelGGL(k_region_fill<float>, grid, blk, 0, st, a);
  else
    hipLaunchKernelGGL(k_region_fill<double>, grid, blk, 0, st, a);
}

void launch_prim_cons(const GridDesc& g, int prec, void* w, const double* bg, double gamma,
                      int to_primitive, hipStream_t st) {
  const long long n = (long long)g.G[0] * g.G[1] * g.G[2];
  const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
  if (g.ndir == 1) {
    if (prec == PREC_MIXED)
      hipLaunchKernelGGL((k_prim_cons<1, float>), grid, blk, 0, st, g, (float*)w, bg, gamma, to_primitive);
    else
      hipLaunchKernelGGL((k_prim_cons<1, double>), grid, blk, 0, st, g, (double*)w, bg, gamma, to_primitive);
  } else if (g.ndir == 2) {
    if (prec == PREC_MIXED)
      hipLaunchKernelGGL((k_prim_cons<2, float>), grid, blk, 0, st, g, (float*)w, bg, gamma, to_primitive);
    else
      hipLaunchKernelGGL((k_prim_cons<2, double>), grid, blk, 0, st, g, (double*)w, bg, gamma, to_primitive);
  } else {
    if (prec == PREC_MIXED)
      hipLaunchKernelGGL((k_prim_cons<3, float>), grid, blk, 0, st, g, (float*)w, bg, gamma, to_primitive);
    else
      hipLaunchKernelGGL((k_prim_cons<3, double>), grid, blk, 0, st, g, (double*)w, bg, gamma, to_primitive);
  }
}

void launch_ghost_store(const GhostArgs& a, void* store, hipStream_t st) {
  const GridDesc& g = a.g;
  long long ncell = (long long)g.G[0] * g.G[1] * g.G[2];
  long long n = ncell - mesh_cells(g);
  if (n <= 0) return;
  const int bs = 256;
  const dim3 grid((unsigned)((n + bs - 1) / bs));
  if (a.prec == PREC_MIXED)
    hipLaunchKernelGGL(k_ghost_store<float>, grid, dim3(bs), 0, st, a, (float*)store);
  else
    hipLaunchKernelGGL(k_ghost_store<double>, grid, dim3(bs), 0, st, a, (double*)store);
}

void launch_zero_u64(unsigned long long* p, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
}

void launch_axpy4(int prec, long long n, void* out, const void* a0, double c0, const void* a1,
                  double c1, const void* a2, double c2, const void* a3, double c3,
                  hipStream_t st) {
  const int bs = 256;
  const dim3 grid((unsigned)((n + bs - 1) / bs));
  if (prec == PREC_MIXED)
    hipLaunchKernelGGL(k_axpy4<float>, grid, dim3(bs), 0, st, n, (float*)out, (const float*)a0,
                       c0, (const float*)a1, c1, (const float*)a2, c2, (const float*)a3, c3);
  else
    hipLaunchKernelGGL(k_axpy4<double>, grid, dim3(bs), 0, st, n, (double*)out,
                       (const double*)a0, c0, (const double*)a1, c1, (const double*)a2, c2,
                       (const double*)a3, c3);
}

void launch_to_float(long long n, float* out, const double* in, hipStream_t st) {
  const int bs = 256;
  hipLaunchKernelGGL((k_convert<float, double>), dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0,
                     st, n, out, in);
}

void launch_to_double(long long n, double* out, const float* in, hipStream_t st) {
  const int bs = 256;
  hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0,
                     st, n, out, in);
}

}  // namespace sac
