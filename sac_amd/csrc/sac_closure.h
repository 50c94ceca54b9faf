// sac_closure.h -- the learned hyperdiffusion closure: launchers of the
// sample, training and inference kernels (sac_closure.hip).
//
// The closure predicts setnu's face ratio md3/md1 (src/vacusr.viscosity.t:
// 413-427) of every variable of an axis from a 7-point stencil of rho',
// v_K and b_(K+1) normalised by the local density, sound speed and
// sqrt(rho) c_s (21 inputs).  Outputs, relative to the axis K:
//   0 rho', 1 thermal energy, 2 v_K, 3-4 v of the other axes (ascending),
//   5-6 b of the other axes (ascending); targets and predictions are the
//   ratio divided by 6 (its upper bound).
// Each of the three 7-point stencils is then scaled to unit maximum (the
// ratio does not depend on a profile's amplitude).  Faces where the shock
// sensor (largest normalised jump of rho' or v_K over the stencil, before
// scaling) reaches the threshold or stays below 1e-3 (flat, or structure
// at the noise level, where the ratio is a ratio of tiny differences),
// faces within 3 cells of the grid edge and states without a positive
// sound speed keep the exact ratio.
#pragma once
#include "sac_common.h"
#include "sac_mlp.h"

namespace sac {
namespace closure {

constexpr int kFeat = mlp::IN;   // BF16 features per sample (21 used)
constexpr int kTarg = 8;         // FP32 targets per sample: 7 ratios/6, weight

// Per-sample training data from one state: every mesh cell (face c+1/2)
// of every axis, written at samples [offset, offset + ndim*Nmesh).
void launch_collect(int ndim, int ndir, const GridDesc& g, const double* w, const double* bg,
                    double gamma, float sensor, uint16_t* feat, float* targ, long long offset,
                    hipStream_t st);

// One training step: `nblocks` wavefronts of 64 samples each, sample j of
// the step = ((base + k) * stride) mod n; per-block gradients and losses.
void launch_train(const uint16_t* feat, const float* targ, long long n, long long base,
                  long long stride, const uint16_t* P16, const float* P32, float* partial,
                  float* loss, int nblocks, float scale, hipStream_t st);

// Sum of the per-block gradients and an Adam step on the FP32 master
// parameters, then their BF16 copy.
void launch_adam(float* P32, uint16_t* P16, float* m1, float* m2, const float* partial, int nblocks,
                 float lr, float beta1, float beta2, float eps, int t, hipStream_t st);

// BF16 copy of FP32 parameters.
void launch_to_bf16(const float* P32, uint16_t* P16, hipStream_t st);

// Inference for axis K on state w: face ratios for the 2*ndim+1 variables
// of setnu on that axis (order rho', thermal energy, v_0.., b_l (l != K)),
// stored at the face's left cell, or -1 where the exact ratio is to be used.
// mode 1: the network; mode 2: the exact ratios through the same path.
void launch_infer(int ndim, int ndir, int axis, const GridDesc& g, const double* w,
                  const double* bg, double gamma, float sensor, const uint16_t* P16,
                  const float* P32, float* RF, int mode, hipStream_t st);

}  // namespace closure
}  // namespace sac
