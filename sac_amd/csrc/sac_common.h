// sac_common.h -- device-side layout and parameter blocks shared by the
// HIP kernels of the SAC perturbation-MHD stage.
//
// Layout: every field is a 3-D array over the grid (mesh + 2 ghost layers
// per active axis, src/vacdef.t:23), index 1 fastest as in the reference's
// w(ixG^T,nw), with the row pitch padded for coalescing.  Indices here are
// 0-based: grid 0..G-1, mesh Mlo..Mhi with Mlo=2, Mhi=G-3.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sac {

constexpr int kMaxDim = 3;
constexpr int kMaxDir = 3;
constexpr int kMaxEvolve = 2 + 2 * kMaxDir;  // rho, m(3), e, b(3)
constexpr int kMaxBg = 2 + kMaxDir;          // eb, rhob, bg(3)
constexpr int kMaxPairs = 21;                // setnu (variable, direction) pairs
constexpr int kRedSlots = 8;                 // reduction values per buffer

// storage precision of the evolving fields
enum PrecCode : int { PREC_FP64 = 0, PREC_MIXED = 1 };

// boundary codes (typeB strings, src/vacgrid.t:257-464)
enum BcCode : int {
  BC_CONT = 1,
  BC_FIXED = 2,
  BC_PERIODIC = 3,
  BC_SYMM = 4,
  BC_SYMM0 = 5,
  BC_ASYMM = 6,
  BC_CONT1 = 7,
  BC_FIXED1 = 8,
  BC_GRAD1 = 9,
  BC_CONTCD4 = 10,
  BC_ZERO = 11
};

struct GridDesc {
  int ndim, ndir;
  int G[3];        // grid extents (1 on inactive axes)
  int Mlo[3];      // first mesh index (2 on active axes, 0 otherwise)
  int Mhi[3];      // last mesh index
  long long st[3]; // element strides per axis
  long long fs;    // field stride (elements)
  int per_lo[3];   // 1 if the lower boundary of axis d is periodic (for var 1)
  int per_hi[3];
};

// reduction slots (as uint64 bit patterns of non-negative doubles), per
// register: [d] = max over the mesh of cmax_d (c_tot of setnu) and
// [ndim+d] = max over the mesh of cmax_d/dx_d (getdt_courant), d<ndim.
// The maxviscoef accumulator is a separate buffer (slot 0).
struct RedBuf {
  unsigned long long v[kRedSlots];
};

}  // namespace sac
