// sac_mlp.h -- a small per-cell MLP on the matrix cores of gfx950: BF16
// operands, FP32 accumulation, one wavefront evaluating 64 samples.
//
// Network: 32 inputs (21 used) -> 64 -> 64 -> 64 (ReLU) -> 16 outputs
// (linear, 7 used).  Parameters, FP32 master copy and BF16 working copy,
// are packed as W1[64][32] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
// W4[16][64] b4[16] (weights row-major [out][in]).
//
// Layout.  Each layer computes Z[unit][sample] = W[unit][k] X[k][sample] with
// v_mfma_f32_16x16x16bf16_1k, units as MFMA rows (M) and samples as columns
// (N).  For this instruction lane l holds A[l%16][4(l/16)+i], B[4(l/16)+i][l%16]
// and D[4(l/16)+i][l%16], i = 0..3.  Activations live in LDS sample-major,
// act[sample][unit] in BF16, so both the B operand of a layer and the store
// of its D tile are 8-byte accesses of 4 consecutive units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sac {
namespace mlp {

constexpr int IN = 32, H = 64, OUT = 16;
constexpr int NIN = 21, NOUT = 7;  // inputs and outputs in use
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * H, OFF_W4 = OFF_B3 + H,
              OFF_B4 = OFF_W4 + OUT * H, NPARAM = OFF_B4 + OUT;
constexpr int WAVE = 64;  // samples per wavefront

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p) {
  return *reinterpret_cast<const bf16x4*>(p);
}
__device__ __forceinline__ void st4(uint16_t* p, f32x4 v) {
  bf16x4 o;
  o[0] = (short)f2bf(v[0]);
  o[1] = (short)f2bf(v[1]);
  o[2] = (short)f2bf(v[2]);
  o[3] = (short)f2bf(v[3]);
  *reinterpret_cast<bf16x4*>(p) = o;
}

// One layer for the wave's 64 samples: Y[s][m] = act(sum_k W[m][k] X[s][k] + b[m]).
// W: BF16 [M][K] (global or LDS), b: FP32 [M], X: LDS [64][K], Y: LDS [64][M].
template <int M, int K, bool RELU>
__device__ __forceinline__ void layer_fwd(const uint16_t* W, const float* b, const uint16_t* X,
                                          uint16_t* Y, int lane) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int mb = 0; mb < M / 16; mb++) {
#pragma unroll
    for (int nb = 0; nb < WAVE / 16; nb++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < K / 16; kb++) {
        const bf16x4 a = ld4(W + (16 * mb + r) * K + 16 * kb + 4 * q);
        const bf16x4 bb = ld4(X + (16 * nb + r) * K + 16 * kb + 4 * q);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
      }
      f32x4 y;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float z = acc[i] + b[16 * mb + 4 * q + i];
        y[i] = RELU ? fmaxf(z, 0.f) : z;
      }
      st4(Y + (16 * nb + r) * M + 16 * mb + 4 * q, y);
    }
  }
}

// The linear output layer, kept in FP32: Y[s][m] = sum_k W[m][k] X[s][k] + b[m].
template <int M, int K>
__device__ __forceinline__ void layer_out(const uint16_t* W, const float* b, const uint16_t* X,
                                          float* Y, int lane) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int mb = 0; mb < M / 16; mb++) {
#pragma unroll
    for (int nb = 0; nb < WAVE / 16; nb++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < K / 16; kb++) {
        const bf16x4 a = ld4(W + (16 * mb + r) * K + 16 * kb + 4 * q);
        const bf16x4 bb = ld4(X + (16 * nb + r) * K + 16 * kb + 4 * q);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; i++) Y[(16 * nb + r) * M + 16 * mb + 4 * q + i] = acc[i] + b[16 * mb + 4 * q + i];
    }
  }
}

// Forward pass of the wave's 64 samples.  x: LDS [64][IN] inputs; h1,h2,h3:
// LDS [64][H] (h3 may alias h1); out: LDS [64][OUT] FP32.  P: BF16
// parameters (weights), Pf: FP32 parameters (biases are read in FP32).
// Barriers between the layers are the caller's (one wave: __syncthreads or a
// wave barrier).
__device__ __forceinline__ void forward(const uint16_t* P, const float* Pf, const uint16_t* x,
                                        uint16_t* h1, uint16_t* h2, uint16_t* h3, float* out,
                                        int lane) {
  layer_fwd<H, IN, true>(P + OFF_W1, Pf + OFF_B1, x, h1, lane);
  __syncthreads();
  layer_fwd<H, H, true>(P + OFF_W2, Pf + OFF_B2, h1, h2, lane);
  __syncthreads();
  layer_fwd<H, H, true>(P + OFF_W3, Pf + OFF_B3, h2, h3, lane);
  __syncthreads();
  layer_out<OUT, H>(P + OFF_W4, Pf + OFF_B4, h3, out, lane);
  __syncthreads();
}

// ---------------------------------------------------------------------
// Inference with the activations in registers.  The D tile (mb, nb) of a
// layer -- lane l holding Z[16mb + 4(l/16) + i][16nb + l%16] -- is laid out
// exactly as the B operand tile (kb = mb, nb) of the next layer needs it, so
// hidden activations go from the accumulators (bias, ReLU, BF16) straight
// into the next layer's operands without LDS.  Only the input (per-lane
// samples to K-major operand) passes through LDS.

// one hidden layer: H[mb][nb] (BF16x4) from the previous layer's X[kb][nb]
template <int M, int K>
__device__ __forceinline__ void reg_layer(const uint16_t* __restrict__ W, const float* __restrict__ b,
                                          const bf16x4 (&X)[K / 16][WAVE / 16],
                                          bf16x4 (&H)[M / 16][WAVE / 16], int lane) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int mb = 0; mb < M / 16; mb++) {
    f32x4 acc[WAVE / 16];
#pragma unroll
    for (int nb = 0; nb < WAVE / 16; nb++) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < K / 16; kb++) {
      const bf16x4 a = ld4(W + (16 * mb + r) * K + 16 * kb + 4 * q);
#pragma unroll
      for (int nb = 0; nb < WAVE / 16; nb++)
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, X[kb][nb], acc[nb], 0, 0, 0);
    }
    float bias[4];
#pragma unroll
    for (int i = 0; i < 4; i++) bias[i] = b[16 * mb + 4 * q + i];
#pragma unroll
    for (int nb = 0; nb < WAVE / 16; nb++) {
      bf16x4 h;
#pragma unroll
      for (int i = 0; i < 4; i++) h[i] = (short)f2bf(fmaxf(acc[nb][i] + bias[i], 0.f));
      H[mb][nb] = h;
    }
  }
}

// Forward pass of the wave's 64 samples from x (LDS, [64][IN] BF16).
// Returns, in lane l, the outputs 4(l/16)+i (i = 0..3) of samples
// 16nb + l%16 (nb = 0..3): out[nb][i].
__device__ __forceinline__ void forward_regs(const uint16_t* __restrict__ P,
                                             const float* __restrict__ Pf, const uint16_t* x,
                                             f32x4 (&out)[WAVE / 16], int lane) {
  const int r = lane & 15, q = lane >> 4;
  bf16x4 X0[IN / 16][WAVE / 16];
#pragma unroll
  for (int kb = 0; kb < IN / 16; kb++)
#pragma unroll
    for (int nb = 0; nb < WAVE / 16; nb++) X0[kb][nb] = ld4(x + (16 * nb + r) * IN + 16 * kb + 4 * q);
  bf16x4 H1[H / 16][WAVE / 16], H2[H / 16][WAVE / 16];
  reg_layer<H, IN>(P + OFF_W1, Pf + OFF_B1, X0, H1, lane);
  reg_layer<H, H>(P + OFF_W2, Pf + OFF_B2, H1, H2, lane);
  reg_layer<H, H>(P + OFF_W3, Pf + OFF_B3, H2, H1, lane);
  // linear output layer (M = OUT = 16: one row tile), FP32
#pragma unroll
  for (int nb = 0; nb < WAVE / 16; nb++) out[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < H / 16; kb++) {
    const bf16x4 a = ld4(P + OFF_W4 + r * H + 16 * kb + 4 * q);
#pragma unroll
    for (int nb = 0; nb < WAVE / 16; nb++)
      out[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, H1[kb][nb], out[nb], 0, 0, 0);
  }
#pragma unroll
  for (int nb = 0; nb < WAVE / 16; nb++)
#pragma unroll
    for (int i = 0; i < 4; i++) out[nb][i] += Pf[OFF_B4 + 4 * q + i];
}

// dX[s][k] = (sum_m W[m][k] dZ[s][m]) * relu'(X[s][k]), for the hidden layers.
// W: BF16 [M][K]; dZ: LDS [64][M]; X: LDS [64][K] (the layer input, whose
// positivity gives relu'); dX: LDS [64][K].
template <int M, int K>
__device__ __forceinline__ void layer_bwd_data(const uint16_t* W, const uint16_t* dZ,
                                               const uint16_t* X, uint16_t* dX, int lane) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int kb = 0; kb < K / 16; kb++) {
#pragma unroll
    for (int nb = 0; nb < WAVE / 16; nb++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int mb = 0; mb < M / 16; mb++) {
        // A = W^T tile: A[k][m] = W[m][k], rows k = 16kb + r, cols m = 16mb + 4q + i
        bf16x4 a;
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = (short)W[(16 * mb + 4 * q + i) * K + 16 * kb + r];
        const bf16x4 bb = ld4(dZ + (16 * nb + r) * M + 16 * mb + 4 * q);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
      }
      // acc[i] = dA[k = 16kb + 4q + i][s = 16nb + r]
      const bf16x4 xv = ld4(X + (16 * nb + r) * K + 16 * kb + 4 * q);
      f32x4 d;
#pragma unroll
      for (int i = 0; i < 4; i++) d[i] = (bf2f((uint16_t)xv[i]) > 0.f) ? acc[i] : 0.f;
      st4(dX + (16 * nb + r) * K + 16 * kb + 4 * q, d);
    }
  }
}

// Weight gradient of one layer over the wave's samples:
// dW[m][k] = sum_s dZ[s][m] X[s][k], db[m] = sum_s dZ[s][m], written to
// G (FP32, packed like the parameters at offsets ow, ob); every element of
// the layer is written exactly once.
template <int M, int K>
__device__ __forceinline__ void layer_bwd_weight(const uint16_t* dZ, const uint16_t* X, float* G,
                                                 int ow, int ob, int lane) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int mb = 0; mb < M / 16; mb++) {
#pragma unroll
    for (int kb = 0; kb < K / 16; kb++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int sb = 0; sb < WAVE / 16; sb++) {
        bf16x4 a, bb;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          a[i] = (short)dZ[(16 * sb + 4 * q + i) * M + 16 * mb + r];  // A[m][s]
          bb[i] = (short)X[(16 * sb + 4 * q + i) * K + 16 * kb + r];  // B[s][k]
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; i++) G[ow + (16 * mb + 4 * q + i) * K + 16 * kb + r] = acc[i];
    }
  }
  // bias: lane m sums its column
  for (int m = lane; m < M; m += WAVE) {
    float s = 0.f;
    for (int t = 0; t < WAVE; t++) s += bf2f(dZ[t * M + m]);
    G[ob + m] = s;
  }
}

}  // namespace mlp
}  // namespace sac
