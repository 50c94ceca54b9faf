// sac_kernels.h -- argument blocks and launchers of the HIP kernels.
#pragma once
#include "sac_common.h"

namespace sac {

// Pointers to evolving fields, R and T have the storage type of the
// precision (double for PREC_FP64, float for PREC_MIXED); the background and
// dx are always double.
struct AxisArgs {
  GridDesc g;
  int prec;
  const void* wct;       // stage input (time-centred state)
  const double* bg;      // background fields
  const double* dx[3];   // per-axis cell widths, indexed by grid index
  void* R;               // rate accumulator (n_evolve fields)
  void* T;               // cross-term factors (2*ndim*(ndim-1) fields)
  double gamma;
  int visc_on;
  const unsigned long long* ctot_src;  // c_tot per direction of wct
  unsigned long long* red_visc;        // maxviscoef accumulator
  int xcd;                             // XCD count for block renumbering (1 = off)
  const float* RF;                     // learned face ratios of axis K, or null (exact)
  int cross_last;                      // last axis: add the cross-term derivatives along it
  int divbfix;                         // Powell's div B sources (ndim > 1)
};

struct FinalArgs {
  GridDesc g;
  int prec;
  const void* wct;
  const double* bg;
  const double* dx[3];
  const void* R;
  const void* T;
  const void* base;      // out = base + qa*R
  void* out;
  // RK4's last stage (w1 not null): out = (base + qa*R) + (w1 + 2*w2 + w3 - 4*base)/3,
  // where w3 is the stage input wct
  const void* w1;
  const void* w2;
  double qa;
  double gamma;
  double grav[3];
  int grav_on[3];
  int visc_on;
  unsigned long long* red_out;  // cmax reductions of out
  int xcd;
  int skip_last;                // cross-term derivatives along the last axis already in R
  int skip_xy;                  // the x-y cross-term derivatives already in R (plane tiles)
};

struct CmaxArgs {
  GridDesc g;
  int prec;
  const void* w;
  const double* bg;
  const double* dx[3];
  double gamma;
  unsigned long long* red_out;
};

struct GhostArgs {
  GridDesc g;
  int prec;            // storage type of f and snap
  void* f;             // nvar fields to fill in place
  const void* snap;    // stored 'fixed' values (null before they exist)
  const int* bc;       // [nvar][6] boundary codes
  int nvar;
  int keep_fixed;      // initial fill with fullgridini: keep file values
  int general;         // some fill has two sources (cont1; fixed1/grad1 before storage)
};

// One region's fill of one field, in the reference's order (iB, then the
// variable), for grids with 'zero' boundaries (launch_region_fill): the
// ghost layers beyond edge `side` of axis D across the whole grid along the
// other axes, from the field's current values.  'zero' takes contCD4's copy
// (the caller converts the register to primitive variables around it).
struct RegionArgs {
  GridDesc g;
  int prec;            // storage type of f
  void* f;             // the field
  int code;            // boundary code
  int D, side;         // axis and edge (0 lower, 1 upper)
  int keep_fixed;      // first fill with fullgridini: keep the file's values
};
void launch_region_fill(const RegionArgs& a, hipStream_t st);
// primitive (to_primitive = 1) or conserve (0) over the whole grid of a
// register (src/vacphys0.t.mhd:5-53)
void launch_prim_cons(const GridDesc& g, int prec, void* w, const double* bg, double gamma,
                      int to_primitive, hipStream_t st);

// resistive sources with a constant eta > 0 (sac_resist.hip)
struct ResistArgs {
  GridDesc g;
  int prec;
  const void* w;         // stage input
  const double* bg;
  const double* dx[3];
  double* J;             // current density scratch, 3 FP64 fields
  void* R;               // rate accumulator
  double eta;
  int compact;           // 1: addsource_res1, 0: addsource_res2
};
void launch_resistive(const ResistArgs& a, hipStream_t st);

// returns 1 if the cross-term derivatives along `axis` were added (requested
// by a.cross_last and done by the line-tile kernel), else 0
int launch_axis(const AxisArgs& a, int axis, hipStream_t st);
// 1 if the line-tile kernels (which carry the Powell sources) are built
int line_tiles_built();
// the x and y work of a 3-D grid in one pass over z-plane tiles (sac_plane.h),
// the x-y cross terms included; R is written fresh
void launch_plane(const AxisArgs& a, hipStream_t st);
void launch_final(const FinalArgs& a, hipStream_t st);
void launch_cmax(const CmaxArgs& a, hipStream_t st);
void launch_ghost(const GhostArgs& a, hipStream_t st);
// stored ghost values of the fixed/fixed1/grad1 boundaries of a.f (after a fill)
void launch_ghost_store(const GhostArgs& a, void* store, hipStream_t st);
void launch_zero_u64(unsigned long long* p, int n, hipStream_t st);
void launch_axpy4(int prec, long long n, void* out, const void* a0, double c0, const void* a1,
                  double c1, const void* a2, double c2, const void* a3, double c3,
                  hipStream_t st);
void launch_to_float(long long n, float* out, const double* in, hipStream_t st);
void launch_to_double(long long n, double* out, const float* in, hipStream_t st);

}  // namespace sac
