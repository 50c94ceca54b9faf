// sac_resist.hip -- resistive MHD sources with a constant eta > 0
// (src/vacphys.mhdres.t), added to the rate accumulator R after the line
// kernels and before the combining kernel.
//
//   k_current: J_i = sum lvc(i,j,k) gradient(b_k + bg_k)_j on the mesh and
//              one layer around it (getcurrent, :48-93), where gradient is
//              the module's own NEGATIVE central difference -(q+ - q-)/dx/2
//              (:304-317);
//   k_resist:  compactres=F (addsource_res2, :223-301):
//                b_i -= sum lvc(i,j,k) gradient(J_k eta)_j
//                e   += sum lvc(i,j,k) gradient((b_j+bg_j) J_k eta)_i
//              compactres=T (addsource_res1, :96-220):
//                b_i += eta lapl(b_i+bg_i),  e += eta lapl(b_i+bg_i)(b_i+bg_i)
//                e   += eta J^2
// The rates are in FP64 and added to R (its storage type) once per cell.
#include "sac_device.h"
#include "sac_kernels.h"

namespace sac {

namespace {

__device__ __forceinline__ int levi(int i, int j, int k) {
  if (i == j || j == k || i == k) return 0;
  return ((i == 1 && j == 2) || (i == 2 && j == 3) || (i == 3 && j == 1)) ? 1 : -1;
}

// first current component getcurrent computes (7 - 2 ndir) and the first one
// with a term (idirmin)
template <int NDIM, int NDIR>
struct Cur {
  static constexpr int kMin0 = 7 - 2 * NDIR;
  static constexpr int kMin = (NDIR == 2) ? 3 : (NDIM >= 2 ? 1 : 2);
};

template <typename S>
__device__ __forceinline__ double btot(const ResistArgs& a, const S* W, int k, long long i) {
  return (double)W[(2 + a.g.ndir + (k - 1)) * a.g.fs + i] + a.bg[(2 + (k - 1)) * a.g.fs + i];
}

template <int NDIM, int NDIR, typename S>
__global__ void __launch_bounds__(256) k_current(ResistArgs a) {
  const GridDesc& g = a.g;
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const bool act = d < NDIM;
    lo[d] = act ? g.Mlo[d] - 1 : g.Mlo[d];
    ext[d] = g.Mhi[d] - g.Mlo[d] + 1 + (act ? 2 : 0);
  }
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ext[0] * ext[1] * ext[2]) return;
  int c[3];
  decode_box(t, lo, ext, c);
  const S* W = (const S*)a.w;
  const long long ic = lin(g, c[0], c[1], c[2]);
#pragma unroll
  for (int idir = Cur<NDIM, NDIR>::kMin0; idir <= 3; idir++) {
    double J = 0.0;
#pragma unroll
    for (int jdir = 1; jdir <= NDIM; jdir++)
#pragma unroll
      for (int kdir = 1; kdir <= NDIR; kdir++) {
        const int l = levi(idir, jdir, kdir);
        if (l == 0) continue;
        const long long s = g.st[jdir - 1];
        const double qp = btot(a, W, kdir, ic + s), qm = btot(a, W, kdir, ic - s);
        const double gr = -(qp - qm) / a.dx[jdir - 1][c[jdir - 1]] / 2.0;
        J = (l == 1) ? J + gr : J - gr;
      }
    a.J[(long long)(idir - 1) * g.fs + ic] = J;
  }
}

template <int NDIM, int NDIR, typename S>
__global__ void __launch_bounds__(256) k_resist(ResistArgs a) {
  constexpr int kMin = Cur<NDIM, NDIR>::kMin;
  const GridDesc& g = a.g;
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = g.Mlo[d];
    ext[d] = g.Mhi[d] - g.Mlo[d] + 1;
  }
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ext[0] * ext[1] * ext[2]) return;
  int c[3];
  decode_box(t, lo, ext, c);
  const S* W = (const S*)a.w;
  S* R = (S*)a.R;
  const long long ic = lin(g, c[0], c[1], c[2]);
  const long long fs = g.fs;
  const double eta = a.eta;
  const int ib = 2 + NDIR, ie = 1 + NDIR;  // R index of b_1 and of e
  auto Jk = [&](int k, long long i) { return a.J[(long long)(k - 1) * fs + i]; };
  double de = 0.0;
  if (!a.compact) {
#pragma unroll
    for (int i = 1; i <= NDIR; i++) {
      double acc = 0.0;
#pragma unroll
      for (int jdir = 1; jdir <= NDIM; jdir++)
#pragma unroll
        for (int kdir = kMin; kdir <= 3; kdir++) {
          const int l = levi(i, jdir, kdir);
          if (l == 0) continue;
          const long long s = g.st[jdir - 1];
          const double tp = Jk(kdir, ic + s) * eta, tm = Jk(kdir, ic - s) * eta;
          const double gr = -(tp - tm) / a.dx[jdir - 1][c[jdir - 1]] / 2.0;
          acc = (l == 1) ? acc - gr : acc + gr;
        }
      const long long ir = (long long)(ib + i - 1) * fs + ic;
      R[ir] = (S)((double)R[ir] + acc);
    }
#pragma unroll
    for (int idir = 1; idir <= NDIM; idir++)
#pragma unroll
      for (int jdir = 1; jdir <= NDIR; jdir++)
#pragma unroll
        for (int kdir = kMin; kdir <= 3; kdir++) {
          const int l = levi(idir, jdir, kdir);
          if (l == 0) continue;
          const long long s = g.st[idir - 1];
          const double tp = btot(a, W, jdir, ic + s) * Jk(kdir, ic + s) * eta;
          const double tm = btot(a, W, jdir, ic - s) * Jk(kdir, ic - s) * eta;
          const double gr = -(tp - tm) / a.dx[idir - 1][c[idir - 1]] / 2.0;
          de = (l == 1) ? de + gr : de - gr;
        }
  } else {
#pragma unroll
    for (int i = 1; i <= NDIR; i++) {
      const double q0 = btot(a, W, i, ic);
      double lap = 0.0;
#pragma unroll
      for (int idim = 1; idim <= NDIM; idim++) {
        const long long s = g.st[idim - 1];
        const double dxd = a.dx[idim - 1][c[idim - 1]];
        lap = lap + (btot(a, W, i, ic + s) - 2 * q0 + btot(a, W, i, ic - s)) / (dxd * dxd);
      }
      lap = lap * eta;
      const long long ir = (long long)(ib + i - 1) * fs + ic;
      R[ir] = (S)((double)R[ir] + lap);
      de += lap * q0;
    }
    double j2 = 0.0;
#pragma unroll
    for (int k = kMin; k <= 3; k++) j2 += Jk(k, ic) * Jk(k, ic);
    de += eta * j2;
  }
  const long long ir = (long long)ie * fs + ic;
  R[ir] = (S)((double)R[ir] + de);
}

template <int NDIM, int NDIR, typename S>
void launch_nd(const ResistArgs& a, hipStream_t st) {
  const GridDesc& g = a.g;
  long long nJ = 1, nM = 1;
  for (int d = 0; d < 3; d++) {
    const long long n = g.Mhi[d] - g.Mlo[d] + 1;
    nM *= n;
    nJ *= n + (d < NDIM ? 2 : 0);
  }
  hipLaunchKernelGGL((k_current<NDIM, NDIR, S>), dim3((unsigned)((nJ + 255) / 256)), dim3(256), 0, st,
                     a);
  hipLaunchKernelGGL((k_resist<NDIM, NDIR, S>), dim3((unsigned)((nM + 255) / 256)), dim3(256), 0, st,
                     a);
}

template <typename S>
void launch_t(const ResistArgs& a, hipStream_t st) {
  const int nd = a.g.ndim, nr = a.g.ndir;
  if (nd == 1 && nr == 2) launch_nd<1, 2, S>(a, st);
  else if (nd == 1 && nr == 3) launch_nd<1, 3, S>(a, st);
  else if (nd == 2 && nr == 2) launch_nd<2, 2, S>(a, st);
  else if (nd == 2 && nr == 3) launch_nd<2, 3, S>(a, st);
  else if (nd == 3 && nr == 3) launch_nd<3, 3, S>(a, st);
}

}  // namespace

void launch_resistive(const ResistArgs& a, hipStream_t st) {
  if (a.prec == PREC_MIXED)
    launch_t<float>(a, st);
  else
    launch_t<double>(a, st);
}

}  // namespace sac
