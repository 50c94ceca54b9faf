// sac_kernels.h -- argument blocks and launchers of the HIP kernels.
#pragma once
#include "sac_common.h"

namespace sac {

struct AxisArgs {
  GridDesc g;
  const double* wct;     // stage input (time-centred state)
  const double* bg;      // background fields
  const double* dx[3];   // per-axis cell widths, indexed by grid index
  double* R;             // rate accumulator (n_evolve fields)
  double* T;             // cross-term factors (2*ndim*(ndim-1) fields)
  double gamma;
  int visc_on;
  const unsigned long long* ctot_src;  // c_tot per direction of wct
  unsigned long long* red_visc;        // maxviscoef accumulator
  int xcd;                             // XCD count for block renumbering (1 = off)
};

struct FinalArgs {
  GridDesc g;
  const double* wct;
  const double* bg;
  const double* dx[3];
  const double* R;
  const double* T;
  const double* base;    // out = base + qa*R
  double* out;
  const double* acc_in;  // acc_out = acc_in + qb*R (acc_out may be null)
  double* acc_out;
  double qa, qb;
  double gamma;
  double grav[3];
  int grav_on[3];
  int visc_on;
  unsigned long long* red_out;  // cmax reductions of out
  int xcd;
};

struct CmaxArgs {
  GridDesc g;
  const double* w;
  const double* bg;
  const double* dx[3];
  double gamma;
  unsigned long long* red_out;
};

struct GhostArgs {
  GridDesc g;
  double* f;           // nvar fields to fill in place
  const double* snap;  // stored 'fixed' values (null before they exist)
  const int* bc;       // [nvar][6] boundary codes
  int nvar;
  int keep_fixed;      // initial fill with fullgridini: keep file values
};

void launch_axis(const AxisArgs& a, int axis, hipStream_t st);
void launch_final(const FinalArgs& a, hipStream_t st);
void launch_cmax(const CmaxArgs& a, hipStream_t st);
void launch_ghost(const GhostArgs& a, hipStream_t st);
void launch_zero_u64(unsigned long long* p, int n, hipStream_t st);
void launch_axpy4(long long n, double* out, const double* a0, double c0, const double* a1,
                  double c1, const double* a2, double c2, const double* a3, double c3,
                  hipStream_t st);

}  // namespace sac
