// sac_closure.hip -- sample generation, training (forward + backward on the
// matrix cores) and inference of the learned hyperdiffusion closure.  See
// sac_closure.h for what the network predicts; the exact quantity it
// replaces is setnu's ratio md3/md1 (src/vacusr.viscosity.t:413-427).
#include "sac_closure.h"

#include "sac_device.h"

namespace sac {
namespace closure {

using mlp::bf2f;
using mlp::f2bf;

// output slot of setnu variable p (k_axis order) on axis K
template <int NDIM, int K>
__device__ __forceinline__ int slot_of(int p) {
  if (p < 2) return p;
  if (p < 2 + NDIM) {
    const int l = p - 2;
    return (l == K) ? 2 : 3 + (l < K ? l : l - 1);
  }
  return 5 + (p - 2 - NDIM);
}

// 21 features of the face between cells c and c+1 along K (7-point stencil
// centred on c) and the shock sensor; false where the exact ratio must be
// used (stencil leaving the grid, no positive sound speed).
template <int NDIM, int NDIR, int K>
__device__ __forceinline__ bool features(const GridDesc& g, const double* __restrict__ W,
                                         const double* __restrict__ B, double gamma,
                                         const int (&c)[3], float (&x)[mlp::IN], float& sensor) {
#pragma unroll
  for (int k = 0; k < mlp::IN; k++) x[k] = 0.f;
  sensor = 1e30f;
  const int cK = c[K];
  if (cK - 3 < 0 || cK + 3 > g.G[K] - 1) return false;
  const long long fs = g.fs, sk = g.st[K];
  const long long ic = lin(g, c[0], c[1], c[2]);
  Pt<NDIR, double> p0;
  load_pt<NDIR, double, double>(W, B, fs, ic, p0);
  const double rt = p0.rho + p0.rhob, inv = 1.0 / rt;
  double sm2 = 0.0, sb2 = 0.0, sbbg = 0.0, sbg2 = 0.0;
#pragma unroll
  for (int d = 0; d < NDIR; d++) {
    sm2 += p0.m[d] * p0.m[d];
    sb2 += p0.b[d] * p0.b[d];
    sbbg += p0.b[d] * p0.bg[d];
    sbg2 += p0.bg[d] * p0.bg[d];
  }
  const double pth = (gamma - 1.0) * (p0.e - (0.5 * sm2 * inv + 0.5 * sb2 + sbbg));
  const double cs2 = gamma * (pth + (gamma - 1.0) * (p0.eb - 0.5 * sbg2)) * inv;
  if (!(cs2 > 0.0)) return false;
  const double cs = sqrt(cs2), bsc = cs * sqrt(rt);
  const double ics = 1.0 / cs, ibsc = 1.0 / bsc;
  constexpr int KB = (K + 1) % NDIR;
  const double v0 = p0.m[K] * inv, b0 = p0.b[KB];
  float smax = 0.f;
#pragma unroll
  for (int s = 0; s < 7; s++) {
    const long long i = ic + (long long)(s - 3) * sk;
    const double rho = W[i];
    const double v = W[(1 + K) * fs + i] / (rho + B[fs + i]);
    const double b = W[(2 + NDIR + KB) * fs + i];
    const float xr = (float)((rho - p0.rho) * inv), xv = (float)((v - v0) * ics);
    x[s] = xr;
    x[7 + s] = xv;
    x[14 + s] = (float)((b - b0) * ibsc);
    smax = fmaxf(smax, fmaxf(fabsf(xr), fabsf(xv)));
  }
  sensor = smax;
  // the ratio md3/md1 of a profile does not depend on its amplitude: each of
  // the three stencils is scaled to unit maximum, the amplitude stays in the
  // sensor
#pragma unroll
  for (int g3 = 0; g3 < 3; g3++) {
    float mx = 0.f;
#pragma unroll
    for (int s = 0; s < 7; s++) mx = fmaxf(mx, fabsf(x[7 * g3 + s]));
    const float sc = (mx > 1e-30f) ? 1.f / mx : 0.f;
#pragma unroll
    for (int s = 0; s < 7; s++) x[7 * g3 + s] *= sc;
  }
  return true;
}

// the network is used for faces with kFlat <= sensor < threshold; below
// kFlat the stencil is flat or carries only noise-level structure (the ratio
// there is a ratio of tiny differences), above the threshold it crosses a
// shock: both keep the exact ratio
constexpr float kFlat = 1e-3f;

// exact face ratios (targets) of the face c+1/2 for the 2*NDIM+1 variables,
// in the q definitions of the exact kernel
template <int NDIM, int NDIR, int K>
__device__ __forceinline__ void exact_ratios(const GridDesc& g, const double* __restrict__ W,
                                             const double* __restrict__ B, const int (&c)[3],
                                             float (&t)[7]) {
#pragma clang fp contract(off)
  const long long fs = g.fs, sk = g.st[K];
  const long long ic = lin(g, c[0], c[1], c[2]);
  constexpr int NP = 2 * NDIM + 1;
  double inv[6];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const long long i = ic + (long long)(s - 2) * sk;
    inv[s] = 1.0 / (W[i] + B[fs + i]);
  }
#pragma unroll
  for (int o = 0; o < 7; o++) t[o] = -1.f;
#pragma unroll
  for (int p = 0; p < NP; p++) {
    double q[6];
    int lb = -1;
    if (p >= 2 + NDIM) {
      const int n = p - 2 - NDIM;
      lb = (n < K) ? n : n + 1;
    }
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const long long i = ic + (long long)(s - 2) * sk;
      if (p == 0) {
        q[s] = W[i];
      } else if (p == 1) {
        double sb = 0.0, sm = 0.0;
#pragma unroll
        for (int cd = 0; cd < NDIR; cd++) {
          const double mc = W[(1 + cd) * fs + i], bc = W[(2 + NDIR + cd) * fs + i];
          sb += bc * bc;
          sm += mc * mc;
        }
        q[s] = W[(1 + NDIR) * fs + i] - 0.5 * (sb + sm * inv[s]);
      } else if (p < 2 + NDIM) {
        q[s] = W[(p - 1) * fs + i] * inv[s];
      } else {
        q[s] = W[(2 + NDIR + lb) * fs + i];
      }
    }
    t[slot_of<NDIM, K>(p)] = (float)(face_ratio<double>(q) / 6.0);
  }
}

template <int NDIM, int NDIR, int K>
__global__ void __launch_bounds__(256) k_collect(GridDesc g, const double* w, const double* bg,
                                                 double gamma, float thr, uint16_t* feat,
                                                 float* targ, long long offset) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = g.Mlo[d];
    ext[d] = g.Mhi[d] - g.Mlo[d] + 1;
  }
  if (t >= (long long)ext[0] * ext[1] * ext[2]) return;
  int c[3];
  decode_box(t, lo, ext, c);
  float x[mlp::IN], s;
  const bool ok = features<NDIM, NDIR, K>(g, w, bg, gamma, c, x, s);
  float tg[7];
#pragma unroll
  for (int o = 0; o < 7; o++) tg[o] = -1.f;
  if (ok) exact_ratios<NDIM, NDIR, K>(g, w, bg, c, tg);  // stencil inside the grid
  const long long j = offset + t;
  uint16_t* fo = feat + j * kFeat;
#pragma unroll
  for (int k = 0; k < kFeat; k++) fo[k] = f2bf(x[k]);
  float* to = targ + j * kTarg;
#pragma unroll
  for (int o = 0; o < 7; o++) to[o] = tg[o];
  to[7] = (ok && s >= kFlat && s < thr) ? 1.f : 0.f;
}

// ---------------------------------------------------------------------
// training: one wavefront (64 samples) per block

__global__ void __launch_bounds__(64) k_train(const uint16_t* feat, const float* targ,
                                              long long n, long long base, long long stride,
                                              const uint16_t* P16, const float* P32,
                                              float* partial, float* loss, float scale) {
  using namespace mlp;
  __shared__ __attribute__((aligned(16))) uint16_t x[WAVE * IN];
  __shared__ __attribute__((aligned(16))) uint16_t h1[WAVE * H], h2[WAVE * H], h3[WAVE * H];
  __shared__ __attribute__((aligned(16))) uint16_t dA[WAVE * H], dB[WAVE * H], d4[WAVE * OUT];
  __shared__ __attribute__((aligned(16))) float out[WAVE * OUT];
  const int lane = threadIdx.x;
  float* G = partial + (size_t)blockIdx.x * NPARAM;
  const unsigned long long j =
      ((unsigned long long)(base + (long long)blockIdx.x * WAVE + lane) *
       (unsigned long long)stride) % (unsigned long long)n;
  const uint16_t* fi = feat + j * kFeat;
#pragma unroll
  for (int k = 0; k < IN; k++) x[lane * IN + k] = fi[k];
  __syncthreads();
  forward(P16, P32, x, h1, h2, h3, out, lane);
  // loss sum_o w (y_o - t_o)^2 over outputs with a target; dL/dy scaled
  const float* ti = targ + j * kTarg;
  const float wgt = ti[7];
  float l2 = 0.f;
#pragma unroll
  for (int o = 0; o < OUT; o++) {
    float d = 0.f;
    if (o < NOUT) {
      const float tv = ti[o];
      if (tv >= 0.f && wgt > 0.f) {
        const float e = out[lane * OUT + o] - tv;
        l2 += wgt * e * e;
        d = 2.f * wgt * e * scale;
      }
    }
    d4[lane * OUT + o] = f2bf(d);
  }
  float ls = l2;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) ls += __shfl_xor(ls, off, 64);
  if (lane == 0) loss[blockIdx.x] = ls;
  __syncthreads();
  layer_bwd_weight<OUT, H>(d4, h3, G, OFF_W4, OFF_B4, lane);
  layer_bwd_data<OUT, H>(P16 + OFF_W4, d4, h3, dA, lane);
  __syncthreads();
  layer_bwd_weight<H, H>(dA, h2, G, OFF_W3, OFF_B3, lane);
  layer_bwd_data<H, H>(P16 + OFF_W3, dA, h2, dB, lane);
  __syncthreads();
  layer_bwd_weight<H, H>(dB, h1, G, OFF_W2, OFF_B2, lane);
  layer_bwd_data<H, H>(P16 + OFF_W2, dB, h1, dA, lane);
  __syncthreads();
  layer_bwd_weight<H, IN>(dA, x, G, OFF_W1, OFF_B1, lane);
}

__global__ void k_adam(float* P32, uint16_t* P16, float* m1, float* m2, const float* partial,
                       int nblocks, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= mlp::NPARAM) return;
  float g = 0.f;
  for (int b = 0; b < nblocks; b++) g += partial[(size_t)b * mlp::NPARAM + k];
  const float m = b1 * m1[k] + (1.f - b1) * g;
  const float v = b2 * m2[k] + (1.f - b2) * g * g;
  m1[k] = m;
  m2[k] = v;
  const float p = P32[k] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  P32[k] = p;
  P16[k] = f2bf(p);
}

__global__ void k_to_bf16(const float* P32, uint16_t* P16) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < mlp::NPARAM) P16[k] = f2bf(P32[k]);
}

// ---------------------------------------------------------------------
// inference: four wavefronts per block, one face per lane; the network runs
// with its activations in registers (mlp::forward_regs), LDS only carries
// the features into the K-major operand layout

// setnu variable of output slot o on axis K (-1: slot unused)
template <int NDIM, int K>
__device__ __forceinline__ int var_of_slot(int o) {
  int v = -1;
#pragma unroll
  for (int p = 0; p < 2 * NDIM + 1; p++)
    if (slot_of<NDIM, K>(p) == o) v = p;
  return v;
}

// mode 1: network ratios; mode 2: the exact ratios, in FP32, through the
// same storage (a check of the plumbing: the run then follows the exact one
// to FP32 rounding of the ratio)
template <int NDIM, int NDIR, int K>
__global__ void __launch_bounds__(256) k_infer(GridDesc g, const double* w, const double* bg,
                                               double gamma, float thr, const uint16_t* P16,
                                               const float* P32, float* RF, int mode) {
  using namespace mlp;
  constexpr int NP = 2 * NDIM + 1;
  __shared__ __attribute__((aligned(16))) uint16_t xs[4][WAVE * IN];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // faces c+1/2 for c = Mlo-1 .. Mhi along K, mesh along the other axes
  int lo[3], ext[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = g.Mlo[d];
    ext[d] = g.Mhi[d] - g.Mlo[d] + 1;
  }
  lo[K] -= 1;
  ext[K] += 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = t < (long long)ext[0] * ext[1] * ext[2];
  int c[3] = {lo[0], lo[1], lo[2]};
  if (active) decode_box(t, lo, ext, c);
  float x[IN], s = 1e30f;
  bool ok = false;
  if (active) ok = features<NDIM, NDIR, K>(g, w, bg, gamma, c, x, s);
  const bool use = ok && s >= kFlat && s < thr;
  const long long ic = lin(g, c[0], c[1], c[2]);
  if (mode == 2) {
    if (active) {
      float ex[7];
      if (use) exact_ratios<NDIM, NDIR, K>(g, w, bg, c, ex);
#pragma unroll
      for (int p = 0; p < NP; p++)
        RF[(long long)p * g.fs + ic] = use ? fminf(fmaxf(6.f * ex[slot_of<NDIM, K>(p)], 0.f), 6.f) : -1.f;
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < IN; k++) xs[wv][lane * IN + k] = use ? f2bf(x[k]) : (uint16_t)0;
  __syncthreads();
  f32x4 out[WAVE / 16];
  forward_regs(P16, P32, xs[wv], out, lane);
  // lane (r, q) holds outputs 4q..4q+3 of samples 16nb + r: each writes those
  // of its samples, with the sample's cell index and flags from its lane
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int nb = 0; nb < WAVE / 16; nb++) {
    const int src = 16 * nb + r;
    const long long ics = __shfl(ic, src, 64);
    const int uses = __shfl((int)use, src, 64);
    const int acts = __shfl((int)active, src, 64);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int p = var_of_slot<NDIM, K>(4 * q + i);
      if (acts && p >= 0)
        RF[(long long)p * g.fs + ics] = uses ? fminf(fmaxf(6.f * out[nb][i], 0.f), 6.f) : -1.f;
    }
  }
}

// ---------------------------------------------------------------------
// launchers

template <int NDIM, int NDIR>
static void collect_nd(const GridDesc& g, const double* w, const double* bg, double gamma,
                       float thr, uint16_t* feat, float* targ, long long offset, hipStream_t st) {
  long long n = 1;
  for (int d = 0; d < 3; d++) n *= g.Mhi[d] - g.Mlo[d] + 1;
  const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
  hipLaunchKernelGGL((k_collect<NDIM, NDIR, 0>), grid, blk, 0, st, g, w, bg, gamma, thr, feat,
                     targ, offset);
  if constexpr (NDIM > 1)
    hipLaunchKernelGGL((k_collect<NDIM, NDIR, 1>), grid, blk, 0, st, g, w, bg, gamma, thr, feat,
                       targ, offset + n);
  if constexpr (NDIM > 2)
    hipLaunchKernelGGL((k_collect<NDIM, NDIR, 2>), grid, blk, 0, st, g, w, bg, gamma, thr, feat,
                       targ, offset + 2 * n);
}

void launch_collect(int ndim, int ndir, const GridDesc& g, const double* w, const double* bg,
                    double gamma, float sensor, uint16_t* feat, float* targ, long long offset,
                    hipStream_t st) {
  if (ndim == 1 && ndir == 1) collect_nd<1, 1>(g, w, bg, gamma, sensor, feat, targ, offset, st);
  else if (ndim == 1 && ndir == 2) collect_nd<1, 2>(g, w, bg, gamma, sensor, feat, targ, offset, st);
  else if (ndim == 1 && ndir == 3) collect_nd<1, 3>(g, w, bg, gamma, sensor, feat, targ, offset, st);
  else if (ndim == 2 && ndir == 2) collect_nd<2, 2>(g, w, bg, gamma, sensor, feat, targ, offset, st);
  else if (ndim == 2 && ndir == 3) collect_nd<2, 3>(g, w, bg, gamma, sensor, feat, targ, offset, st);
  else if (ndim == 3 && ndir == 3) collect_nd<3, 3>(g, w, bg, gamma, sensor, feat, targ, offset, st);
}

void launch_train(const uint16_t* feat, const float* targ, long long n, long long base,
                  long long stride, const uint16_t* P16, const float* P32, float* partial,
                  float* loss, int nblocks, float scale, hipStream_t st) {
  hipLaunchKernelGGL(k_train, dim3(nblocks), dim3(64), 0, st, feat, targ, n, base, stride, P16,
                     P32, partial, loss, scale);
}

void launch_adam(float* P32, uint16_t* P16, float* m1, float* m2, const float* partial, int nblocks,
                 float lr, float beta1, float beta2, float eps, int t, hipStream_t st) {
  const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3((mlp::NPARAM + 255) / 256), dim3(256), 0, st, P32, P16, m1, m2,
                     partial, nblocks, lr, beta1, beta2, eps, c1, c2);
}

void launch_to_bf16(const float* P32, uint16_t* P16, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, dim3((mlp::NPARAM + 255) / 256), dim3(256), 0, st, P32, P16);
}

template <int NDIM, int NDIR>
static void infer_nd(int axis, const GridDesc& g, const double* w, const double* bg, double gamma,
                     float thr, const uint16_t* P16, const float* P32, float* RF, int mode,
                     hipStream_t st) {
  long long n = 1;
  for (int d = 0; d < 3; d++) n *= g.Mhi[d] - g.Mlo[d] + 1 + (d == axis ? 1 : 0);
  const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
  if (axis == 0)
    hipLaunchKernelGGL((k_infer<NDIM, NDIR, 0>), grid, blk, 0, st, g, w, bg, gamma, thr, P16, P32,
                       RF, mode);
  if constexpr (NDIM > 1)
    if (axis == 1)
      hipLaunchKernelGGL((k_infer<NDIM, NDIR, 1>), grid, blk, 0, st, g, w, bg, gamma, thr, P16,
                         P32, RF, mode);
  if constexpr (NDIM > 2)
    if (axis == 2)
      hipLaunchKernelGGL((k_infer<NDIM, NDIR, 2>), grid, blk, 0, st, g, w, bg, gamma, thr, P16,
                         P32, RF, mode);
}

void launch_infer(int ndim, int ndir, int axis, const GridDesc& g, const double* w,
                  const double* bg, double gamma, float sensor, const uint16_t* P16,
                  const float* P32, float* RF, int mode, hipStream_t st) {
  if (ndim == 1 && ndir == 1) infer_nd<1, 1>(axis, g, w, bg, gamma, sensor, P16, P32, RF, mode, st);
  else if (ndim == 1 && ndir == 2) infer_nd<1, 2>(axis, g, w, bg, gamma, sensor, P16, P32, RF, mode, st);
  else if (ndim == 1 && ndir == 3) infer_nd<1, 3>(axis, g, w, bg, gamma, sensor, P16, P32, RF, mode, st);
  else if (ndim == 2 && ndir == 2) infer_nd<2, 2>(axis, g, w, bg, gamma, sensor, P16, P32, RF, mode, st);
  else if (ndim == 2 && ndir == 3) infer_nd<2, 3>(axis, g, w, bg, gamma, sensor, P16, P32, RF, mode, st);
  else if (ndim == 3 && ndir == 3) infer_nd<3, 3>(axis, g, w, bg, gamma, sensor, P16, P32, RF, mode, st);
}

}  // namespace closure
}  // namespace sac
