// sac_abi.hip -- C ABI (include/sac_amd.h) over the HIP kernels: device
// buffers, the Runge-Kutta stage sequence, ghost fills and time-step
// reductions.  One handle = one problem on one device; all work is issued
// on the handle's own stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sac_amd.h"
#include "sac_closure.h"
#include "sac_kernels.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) return fail(SAC_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr int kSlots = 4;
constexpr double kSmallDouble = 1.0e-99;  // src/vacdef.t:50
constexpr double kBigDouble = 1.0e99;

// learned-closure state of a handle (sac_closure_*)
struct Closure {
  float* P32 = nullptr;      // FP32 master parameters
  uint16_t* P16 = nullptr;   // BF16 working copy
  float* m1 = nullptr;       // Adam moments
  float* m2 = nullptr;
  float* partial = nullptr;  // per-block gradients [max_blocks][NPARAM]
  float* loss = nullptr;     // per-block loss sums
  int max_blocks = 0;
  float* RF[3] = {nullptr, nullptr, nullptr};  // learned face ratios per axis
  uint16_t* feat = nullptr;  // training samples
  float* targ = nullptr;
  long long cap = 0, n = 0;
  int adam_t = 0;
  int on = 0;
  float sensor = 0.1f;
};

struct TimedLaunch {
  int kernel;
  int kernel2;  // per-axis id of a k_axis launch (-1 otherwise)
  hipEvent_t a, b;
};

struct Dev {
  int device = 0;
  hipStream_t stream = nullptr;
  sac::GridDesc g{};
  int ndim = 0, ndir = 0, nw = 0, ne = 0, nbg = 0, npairs = 0;
  int nx[3] = {1, 1, 1};
  long long ncell_grid = 0;  // G1*G2*G3 (unpadded, host layout)
  int typeadvance = SAC_ADV_FOURSTEP;
  int prec = SAC_PREC_FP64;
  size_t esz = sizeof(double);  // bytes per element of the evolving fields
  int fullgridini = 0;
  double gamma = 0.0;
  double grav[3] = {0, 0, 0};
  int grav_on[3] = {0, 0, 0};
  int visc_on = 0;
  int divbfix = 0;  // Powell's div B sources (addsource_divb), ndim > 1
  // resistive MHD (src/vacphys.mhdres.t): constant eta = eqpar(eta_) > 0,
  // addsource_res1 (compactres) or res2, and the common dtdiffpar (the par
  // value until getdt_visc assigns 0.25)
  double eta = 0.0;
  int compactres = 0;
  double dtdiffpar = 0.5;
  double* J = nullptr;  // current density scratch (3 fields)
  double dtgrav = kBigDouble;
  double dxmin[3] = {0, 0, 0};
  bool any_fixed = false;
  // boundary fills with two sources: cont1 always, fixed1/grad1 until
  // their values are stored (evolving fields, background)
  bool cont1_e = false, ext1_e = false, cont1_b = false, ext1_b = false;
  bool initialized = false;  // getboundary's 'initialized' (src/vacgrid.t:209)
  bool any_zero = false;     // some 'zero' boundary: region-by-region fills
  std::vector<int> bce_h, bcb_h;  // boundary codes [var][6] (evolving, background)
  // an upload after the first getboundary replaced the background's ghost
  // cells with the caller's: the next getboundary fills them again
  bool bg_ghosts_stale = false;
  // device buffers (evolving fields, R, T and snap in the storage precision)
  void* reg[kSlots] = {nullptr, nullptr, nullptr, nullptr};
  double* bg = nullptr;
  void* rate = nullptr;     // R accumulator
  void* cross = nullptr;    // cross-term factors T
  void* snap = nullptr;     // 'fixed' values of the evolving fields
  double* dx[3] = {nullptr, nullptr, nullptr};
  int* bc_e = nullptr;      // [ne][6]
  int* bc_b = nullptr;      // [nbg][6]
  unsigned long long* red = nullptr;  // [kSlots+1][kRedSlots]
  bool red_valid[kSlots] = {false, false, false, false};
  std::vector<double> stage;  // host staging
  long long bytes = 0;
  bool timing = false;
  std::vector<TimedLaunch> timed;
  double t_ms[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  long long t_n[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  // 3-D stages: the x and y work in one pass over z-plane tiles (k_plane)
  // instead of one line-tile pass per axis
  int plane = 1;
  int xcd = 8;  // XCDs of the device (block renumbering of the line kernels)
  Closure* cl = nullptr;
  // asynchronous saves (sac_dev_save_*): a device copy of W taken on the
  // handle's stream, copied to pinned host memory on a second stream (both
  // at the default priority) while later steps run
  void* save_dev = nullptr;      // ne fields, storage precision
  void* save_host = nullptr;     // pinned, ne * ncell_grid elements (host layout)
  hipStream_t save_stream = nullptr;
  hipEvent_t save_taken = nullptr, save_done = nullptr;
  bool save_pending = false;
  std::vector<double> bg_host;   // background of the last upload (static)
};

long long fieldsz(const Dev& d) { return d.g.fs; }

int balloc(Dev& d, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(SAC_E_NOMEM, "hipMalloc of %zu bytes failed: %s", bytes,
                                   hipGetErrorString(e));
  d.bytes += (long long)bytes;
  e = hipMemset(*p, 0, bytes);
  if (e != hipSuccess) return fail(SAC_E_HIP, "hipMemset: %s", hipGetErrorString(e));
  return 0;
}
int dalloc(Dev& d, double** p, long long nelem) {
  return balloc(d, (void**)p, sizeof(double) * (size_t)nelem);
}
// nelem elements of the storage precision
int ealloc(Dev& d, void** p, long long nelem) { return balloc(d, p, d.esz * (size_t)nelem); }

unsigned long long* red_of(Dev& d, int slot) { return d.red + (size_t)slot * sac::kRedSlots; }
unsigned long long* red_visc(Dev& d) { return d.red + (size_t)kSlots * sac::kRedSlots; }

// event helpers
void t_begin(Dev& d, int k, hipEvent_t* a) {
  if (!d.timing) return;
  hipEventCreate(a);
  hipEventRecord(*a, d.stream);
  (void)k;
}
void t_end(Dev& d, int k, hipEvent_t a, int k2 = -1) {
  if (!d.timing) return;
  hipEvent_t b;
  hipEventCreate(&b);
  hipEventRecord(b, d.stream);
  d.timed.push_back({k, k2, a, b});
}
void t_collect(Dev& d) {
  for (auto& t : d.timed) {
    hipEventSynchronize(t.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, t.a, t.b);
    d.t_ms[t.kernel] += ms;
    d.t_n[t.kernel] += 1;
    if (t.kernel2 >= 0) {
      d.t_ms[t.kernel2] += ms;
      d.t_n[t.kernel2] += 1;
    }
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  d.timed.clear();
}

void ghost_fill_evolving(Dev& d, int slot) {
  sac::GhostArgs a{};
  a.g = d.g;
  a.prec = d.prec;
  a.f = d.reg[slot];
  a.bc = d.bc_e;
  a.nvar = d.ne;
  a.snap = d.initialized ? d.snap : nullptr;
  a.keep_fixed = (!d.initialized && d.fullgridini) ? 1 : 0;
  a.general = d.cont1_e || (d.ext1_e && !a.snap && !a.keep_fixed);
  hipEvent_t e0{};
  t_begin(d, SAC_K_GHOST, &e0);
  sac::launch_ghost(a, d.stream);
  t_end(d, SAC_K_GHOST, e0);
}

// getboundary region by region, variable by variable, as the reference
// does (src/vacgrid.t:195-512): the path of grids with 'zero' boundaries,
// whose fills convert the whole register to primitive variables and back.
// The background's own fills are made when its ghosts are not yet filled;
// afterwards only its 'zero' fills recur (for their conversions of the
// register).  The conversions move mesh cells by rounding, so the register's
// reductions are taken again before their next use.
void ghost_fill_sequential(Dev& d, int slot, bool fill_bg) {
  for (int iB = 0; iB < 2 * d.ndim; iB++) {
    for (int iw = 0; iw < d.ne + d.nbg; iw++) {
      const bool ev = iw < d.ne;
      const int code = ev ? d.bce_h[(size_t)iw * 6 + iB] : d.bcb_h[(size_t)(iw - d.ne) * 6 + iB];
      if (!ev && !fill_bg && code != SAC_BC_ZERO) continue;
      sac::RegionArgs r{};
      r.g = d.g;
      r.prec = ev ? d.prec : sac::PREC_FP64;
      r.f = ev ? (void*)((char*)d.reg[slot] + d.esz * (size_t)iw * d.g.fs)
               : (void*)(d.bg + (size_t)(iw - d.ne) * d.g.fs);
      r.code = code;
      r.D = iB / 2;
      r.side = iB % 2;
      r.keep_fixed = (!d.initialized && d.fullgridini) ? 1 : 0;
      const bool zero = code == SAC_BC_ZERO;
      if (zero) sac::launch_prim_cons(d.g, d.prec, d.reg[slot], d.bg, d.gamma, 1, d.stream);
      if (!(zero && !ev && !fill_bg)) sac::launch_region_fill(r, d.stream);
      if (zero) sac::launch_prim_cons(d.g, d.prec, d.reg[slot], d.bg, d.gamma, 0, d.stream);
    }
  }
  d.red_valid[slot] = false;
}

void cmax_reduce(Dev& d, int slot) {
  sac::launch_zero_u64(red_of(d, slot), sac::kRedSlots, d.stream);
  sac::CmaxArgs a{};
  a.g = d.g;
  a.prec = d.prec;
  a.w = d.reg[slot];
  a.bg = d.bg;
  for (int k = 0; k < 3; k++) a.dx[k] = d.dx[k];
  a.gamma = d.gamma;
  a.red_out = red_of(d, slot);
  hipEvent_t e0{};
  t_begin(d, SAC_K_CMAX, &e0);
  sac::launch_cmax(a, d.stream);
  t_end(d, SAC_K_CMAX, e0);
  d.red_valid[slot] = true;
}

// one stage: out = base + qa*R(wct), ghosts of out.  rk4_last: RK4's last
// stage (wct is S3), out also gets (S1 + 2*S2 + S3 - 4*base)/3 of the three
// earlier stage registers (the reference's w1, w2, w3)
void stage(Dev& d, int wct, int base, int out, double qa, bool rk4_last = false) {
  if (!d.red_valid[wct]) cmax_reduce(d, wct);
  sac::AxisArgs x{};
  x.g = d.g;
  x.prec = d.prec;
  x.wct = d.reg[wct];
  x.bg = d.bg;
  for (int k = 0; k < 3; k++) x.dx[k] = d.dx[k];
  x.R = d.rate;
  x.T = d.cross;
  x.gamma = d.gamma;
  x.visc_on = d.visc_on;
  x.ctot_src = red_of(d, wct);
  x.red_visc = red_visc(d);
  const bool learned = d.cl && d.cl->on;
  int skip_last = 0;
  // plane tiles: the x and y work (and their cross terms) in one launch,
  // then the z line tiles (which add the z cross terms), then k_final
  const bool plane = d.plane && d.ndim == 3 && !learned && !d.divbfix && sac::line_tiles_built();
  if (plane) {
    x.xcd = d.xcd;
    x.RF = nullptr;
    x.cross_last = 0;
    x.divbfix = 0;
    hipEvent_t e0{};
    t_begin(d, SAC_K_PLANE, &e0);
    sac::launch_plane(x, d.stream);
    t_end(d, SAC_K_PLANE, e0);
  }
  for (int k = plane ? 2 : 0; k < d.ndim; k++) {
    // axis 1 (rows of a plane): neighbour rows in one XCD's L2.  Axes 0 and 2
    // keep the hardware order, whose round-robin already places the
    // neighbour planes of a row in the same XCD.
    x.xcd = (k == 1) ? d.xcd : 1;
    x.RF = nullptr;
    if (learned) {
      hipEvent_t ei{};
      t_begin(d, SAC_K_CLOSURE, &ei);
      sac::closure::launch_infer(d.ndim, d.ndir, k, d.g, (const double*)d.reg[wct], d.bg, d.gamma,
                                 d.cl->sensor, d.cl->P16, d.cl->P32, d.cl->RF[k], d.cl->on,
                                 d.stream);
      t_end(d, SAC_K_CLOSURE, ei);
      x.RF = d.cl->RF[k];
    }
    x.cross_last = (k == d.ndim - 1) ? 1 : 0;
    x.divbfix = d.divbfix;
    hipEvent_t e0{};
    t_begin(d, SAC_K_AXIS, &e0);
    skip_last = sac::launch_axis(x, k, d.stream);
    t_end(d, SAC_K_AXIS, e0, 5 + k);
  }
  if (d.eta > 0.0) {  // resistive sources into R (addsource, src/vacphys.mhd0.t:240-249)
    sac::ResistArgs ra{};
    ra.g = d.g;
    ra.prec = d.prec;
    ra.w = d.reg[wct];
    ra.bg = d.bg;
    for (int k = 0; k < 3; k++) ra.dx[k] = d.dx[k];
    ra.J = d.J;
    ra.R = d.rate;
    ra.eta = d.eta;
    ra.compact = d.compactres;
    sac::launch_resistive(ra, d.stream);
  }
  sac::launch_zero_u64(red_of(d, out), sac::kRedSlots, d.stream);
  sac::FinalArgs s{};
  s.g = d.g;
  s.prec = d.prec;
  s.wct = d.reg[wct];
  s.bg = d.bg;
  for (int k = 0; k < 3; k++) s.dx[k] = d.dx[k];
  s.R = d.rate;
  s.T = d.cross;
  s.base = d.reg[base];
  s.out = d.reg[out];
  if (rk4_last) {
    s.w1 = d.reg[SAC_SLOT_S1];
    s.w2 = d.reg[SAC_SLOT_S2];
  }
  s.qa = qa;
  s.gamma = d.gamma;
  for (int k = 0; k < 3; k++) {
    s.grav[k] = d.grav[k];
    s.grav_on[k] = d.grav_on[k];
  }
  s.visc_on = d.visc_on;
  s.red_out = red_of(d, out);
  s.xcd = 1;
  s.skip_last = skip_last;
  s.skip_xy = plane ? 1 : 0;
  hipEvent_t e0{};
  t_begin(d, SAC_K_FINAL, &e0);
  sac::launch_final(s, d.stream);
  t_end(d, SAC_K_FINAL, e0);
  d.red_valid[out] = true;
  if (d.any_zero)
    ghost_fill_sequential(d, out, false);
  else
    ghost_fill_evolving(d, out);
}

int check_handle(sac_handle h) {
  if (!h) return fail(SAC_E_ARG, "null handle");
  return 0;
}

// host (unpadded, FP64) <-> device (padded rows, element size esz) copies of
// nfield fields; FP32 device fields are converted on the host, one field at
// a time
int copy_in(Dev& d, void* dst, const double* src, int nfield, size_t esz) {
  const size_t w = esz * (size_t)d.g.G[0];
  const size_t spitch = w;
  const size_t dpitch = esz * (size_t)d.g.st[1];
  const size_t rows = (size_t)d.g.G[1] * (size_t)d.g.G[2];
  std::vector<float> tmp;
  for (int f = 0; f < nfield; f++) {
    const double* s = src + (size_t)f * d.ncell_grid;
    char* dp = (char*)dst + esz * (size_t)f * d.g.fs;
    if (esz == sizeof(double)) {
      HIPCHK(hipMemcpy2DAsync(dp, dpitch, s, spitch, w, rows, hipMemcpyHostToDevice, d.stream));
    } else {
      tmp.resize((size_t)d.ncell_grid);
      for (long long i = 0; i < d.ncell_grid; i++) tmp[i] = (float)s[i];
      HIPCHK(hipMemcpy2DAsync(dp, dpitch, tmp.data(), spitch, w, rows, hipMemcpyHostToDevice,
                              d.stream));
      HIPCHK(hipStreamSynchronize(d.stream));
    }
  }
  HIPCHK(hipStreamSynchronize(d.stream));
  return 0;
}

int copy_out(Dev& d, double* dst, const void* src, int nfield, size_t esz) {
  const size_t w = esz * (size_t)d.g.G[0];
  const size_t dpitch = w;
  const size_t spitch = esz * (size_t)d.g.st[1];
  const size_t rows = (size_t)d.g.G[1] * (size_t)d.g.G[2];
  std::vector<float> tmp;
  for (int f = 0; f < nfield; f++) {
    double* o = dst + (size_t)f * d.ncell_grid;
    const char* sp = (const char*)src + esz * (size_t)f * d.g.fs;
    if (esz == sizeof(double)) {
      HIPCHK(hipMemcpy2DAsync(o, dpitch, sp, spitch, w, rows, hipMemcpyDeviceToHost, d.stream));
    } else {
      tmp.resize((size_t)d.ncell_grid);
      HIPCHK(hipMemcpy2DAsync(tmp.data(), dpitch, sp, spitch, w, rows, hipMemcpyDeviceToHost,
                              d.stream));
      HIPCHK(hipStreamSynchronize(d.stream));
      for (long long i = 0; i < d.ncell_grid; i++) o[i] = (double)tmp[i];
    }
  }
  HIPCHK(hipStreamSynchronize(d.stream));
  return 0;
}

// the reductions [k0, k0+n) of register `slot` on the host (one copy, one wait)
void red_host(Dev& d, int slot, int k0, int n, double* x) {
  unsigned long long v[sac::kRedSlots] = {};
  hipMemcpyAsync(v, red_of(d, slot) + k0, sizeof(v[0]) * n, hipMemcpyDeviceToHost, d.stream);
  hipStreamSynchronize(d.stream);
  memcpy(x, v, sizeof(double) * n);
}

}  // namespace

namespace sac {
// the message of sac_dev_last_error() for entry points in other files
int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
}  // namespace sac

extern "C" {

const char* sac_dev_last_error(void) { return g_err.c_str(); }

int sac_dev_init(int ndim, int ndir, const int* nx, int nw, const int* iws, int niws,
                 const double* eqpar, int neqpar, int grav0_, int nu_, const double* dx,
                 const int* typeB, int fullgridini, int typeadvance, int precision,
                 sac_handle* out) {
  g_err.clear();
  if (!out) return fail(SAC_E_ARG, "out is null");
  *out = nullptr;
  if (ndim < 1 || ndim > 3 || ndir < ndim || ndir > 3)
    return fail(SAC_E_ARG, "need 1<=ndim<=ndir<=3 (ndim=%d ndir=%d)", ndim, ndir);
  if (ndim == 2 && ndir == 2) {
  } else if (!(ndim == 1 || (ndim == 2 && ndir == 3) || (ndim == 3 && ndir == 3))) {
    return fail(SAC_E_UNSUPPORTED, "ndim=%d ndir=%d not built", ndim, ndir);
  }
  if (ndim == 1 && ndir < 1) return fail(SAC_E_ARG, "ndir");
  if (!nx || !eqpar || !dx || !typeB || !iws) return fail(SAC_E_ARG, "null array argument");
  if (nw != 4 + 3 * ndir) return fail(SAC_E_ARG, "nw=%d, expected %d", nw, 4 + 3 * ndir);
  const int ne = 2 + 2 * ndir;
  if (niws != ne) return fail(SAC_E_UNSUPPORTED, "iws must list the %d evolving variables", ne);
  for (int i = 0; i < niws; i++)
    if (iws[i] != i + 1) return fail(SAC_E_UNSUPPORTED, "iws must be 1..%d in order", ne);
  if (neqpar < 2) return fail(SAC_E_ARG, "neqpar=%d < 2", neqpar);
  if (eqpar[0] <= 0.0) return fail(SAC_E_ARG, "gamma must be > 0");
  if (eqpar[1] < 0.0)
    return fail(SAC_E_UNSUPPORTED,
                "resistive MHD with eqpar(eta_)<0 (the user's specialeta) is not built");
  if (eqpar[1] > 0.0 && ndir < 2)
    return fail(SAC_E_ARG, "resistive MHD needs ndir >= 2");
  if (typeadvance < SAC_ADV_ONESTEP || typeadvance > SAC_ADV_JAMESON)
    return fail(SAC_E_UNSUPPORTED, "typeadvance code %d not built", typeadvance);
  if (precision != SAC_PREC_FP64 && precision != SAC_PREC_MIXED)
    return fail(SAC_E_UNSUPPORTED, "precision %d not built", precision);
  for (int d = 0; d < ndim; d++)
    if (nx[d] < 6) return fail(SAC_E_ARG, "nx[%d]=%d < 6", d, nx[d]);
  for (int iw = 0; iw < nw; iw++)
    for (int iB = 0; iB < 2 * ndim; iB++) {
      int c = typeB[iw * 2 * ndim + iB];
      if (c < SAC_BC_CONT || c > SAC_BC_ZERO)
        return fail(SAC_E_UNSUPPORTED, "boundary code %d (var %d, region %d)", c, iw + 1, iB + 1);
    }
  {
    // 'zero' fills run region by region (ghost_fill_sequential), which
    // restores no stored values: not with fixed, fixed1 or grad1
    bool zero = false, stored = false;
    for (int i = 0; i < nw * 2 * ndim; i++) {
      zero |= typeB[i] == SAC_BC_ZERO;
      stored |= typeB[i] == SAC_BC_FIXED || typeB[i] == SAC_BC_FIXED1 || typeB[i] == SAC_BC_GRAD1;
    }
    if (zero && stored)
      return fail(SAC_E_UNSUPPORTED,
                  "'zero' boundaries together with fixed, fixed1 or grad1 are not built");
  }
  // periodic pairing must be consistent (src/vacgrid.t:109-120)
  for (int d = 0; d < ndim; d++) {
    int lo = typeB[0 * 2 * ndim + 2 * d], hi = typeB[0 * 2 * ndim + 2 * d + 1];
    if ((lo == SAC_BC_PERIODIC) != (hi == SAC_BC_PERIODIC))
      return fail(SAC_E_ARG, "Error in BoundSetup: No periodic pair (axis %d)", d + 1);
  }

  int ndev = 0;
  hipError_t he = hipGetDeviceCount(&ndev);
  if (he != hipSuccess || ndev < 1)
    return fail(SAC_E_HIP, "no HIP device available (%s)", hipGetErrorString(he));

  Dev* d = new Dev();
  HIPCHK(hipGetDevice(&d->device));
  d->ndim = ndim;
  d->ndir = ndir;
  d->nw = nw;
  d->ne = ne;
  d->nbg = 2 + ndir;
  d->npairs = ndim * (2 * ndim + 1);
  d->typeadvance = typeadvance;
  d->prec = precision;
  d->esz = (precision == SAC_PREC_MIXED) ? sizeof(float) : sizeof(double);
  d->fullgridini = fullgridini;
  d->gamma = eqpar[0];
  sac::GridDesc& g = d->g;
  g.ndim = ndim;
  g.ndir = ndir;
  long long ncg = 1;
  for (int k = 0; k < 3; k++) {
    if (k < ndim) {
      d->nx[k] = nx[k];
      g.G[k] = nx[k] + 4;
      g.Mlo[k] = 2;
      g.Mhi[k] = nx[k] + 1;
    } else {
      d->nx[k] = 1;
      g.G[k] = 1;
      g.Mlo[k] = 0;
      g.Mhi[k] = 0;
    }
    ncg *= g.G[k];
  }
  d->ncell_grid = ncg;
  const int pitch = (g.G[0] + 15) / 16 * 16;  // 128-byte rows
  g.st[0] = 1;
  g.st[1] = pitch;
  g.st[2] = (long long)pitch * g.G[1];
  g.fs = (long long)pitch * g.G[1] * g.G[2];
  for (int k = 0; k < ndim; k++) {
    g.per_lo[k] = typeB[0 * 2 * ndim + 2 * k] == SAC_BC_PERIODIC;
    g.per_hi[k] = typeB[0 * 2 * ndim + 2 * k + 1] == SAC_BC_PERIODIC;
  }
  // user parameters: gravity at grav0_+d, switch at nu_ (1-based)
  for (int k = 0; k < ndim; k++) {
    int slot = grav0_ + k + 1;
    double gk = (slot >= 1 && slot <= neqpar) ? eqpar[slot - 1] : 0.0;
    d->grav[k] = gk;
    d->grav_on[k] = std::fabs(gk) > kSmallDouble;
  }
  double nu = (nu_ >= 1 && nu_ <= neqpar) ? eqpar[nu_ - 1] : 0.0;
  d->visc_on = std::fabs(nu) > kSmallDouble;
  d->eta = eqpar[1];
  for (int iw = 0; iw < nw; iw++)
    for (int iB = 0; iB < 2 * ndim; iB++)
      if (typeB[iw * 2 * ndim + iB] == SAC_BC_FIXED || typeB[iw * 2 * ndim + iB] == SAC_BC_FIXED1 ||
          typeB[iw * 2 * ndim + iB] == SAC_BC_GRAD1)
        d->any_fixed = true;

  int rc = 0;
  auto bail = [&](int code) {
    sac_dev_free(d);
    return code;
  };
  if (hipStreamCreate(&d->stream) != hipSuccess) return bail(fail(SAC_E_HIP, "hipStreamCreate"));
  for (int s = 0; s < kSlots; s++)
    if ((rc = ealloc(*d, &d->reg[s], (long long)ne * g.fs))) return bail(rc);
  if ((rc = dalloc(*d, &d->bg, (long long)d->nbg * g.fs))) return bail(rc);
  if ((rc = ealloc(*d, &d->rate, (long long)ne * g.fs))) return bail(rc);
  if (d->eta > 0.0 && (rc = dalloc(*d, &d->J, 3LL * g.fs))) return bail(rc);
  if (d->visc_on && ndim > 1 &&
      (rc = ealloc(*d, &d->cross, (long long)2 * ndim * (ndim - 1) * g.fs)))
    return bail(rc);
  if (d->any_fixed) {
    if ((rc = ealloc(*d, &d->snap, (long long)ne * g.fs))) return bail(rc);
  }
  // dx per axis, dxmin per axis over the mesh
  long long off = 0;
  for (int k = 0; k < 3; k++) {
    int n = g.G[k];
    if ((rc = dalloc(*d, &d->dx[k], n))) return bail(rc);
    if (k < ndim) {
      if (hipMemcpy(d->dx[k], dx + off, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SAC_E_HIP, "dx upload"));
      double mn = kBigDouble;
      for (int i = g.Mlo[k]; i <= g.Mhi[k]; i++) {
        if (!(dx[off + i] > 0.0)) return bail(fail(SAC_E_ARG, "dx must be > 0 on the mesh"));
        mn = std::fmin(mn, dx[off + i]);
      }
      d->dxmin[k] = mn;
      off += n;
    } else {
      double one = 1.0;
      hipMemcpy(d->dx[k], &one, sizeof(double), hipMemcpyHostToDevice);
    }
  }
  // boundary codes per register
  std::vector<int>& bce = d->bce_h;
  std::vector<int>& bcb = d->bcb_h;
  bce.assign((size_t)ne * 6, SAC_BC_CONT);
  bcb.assign((size_t)d->nbg * 6, SAC_BC_CONT);
  for (int iw = 0; iw < nw; iw++)
    for (int iB = 0; iB < 2 * ndim; iB++) {
      int c = typeB[iw * 2 * ndim + iB];
      const bool c1 = c == SAC_BC_CONT1, x1 = c == SAC_BC_FIXED1 || c == SAC_BC_GRAD1;
      d->any_zero |= c == SAC_BC_ZERO;
      if (iw < ne) {
        bce[(size_t)iw * 6 + iB] = c;
        d->cont1_e |= c1;
        d->ext1_e |= x1;
      } else {
        bcb[(size_t)(iw - ne) * 6 + iB] = c;
        d->cont1_b |= c1;
        d->ext1_b |= x1;
      }
    }
  if (hipMalloc((void**)&d->bc_e, sizeof(int) * bce.size()) != hipSuccess ||
      hipMalloc((void**)&d->bc_b, sizeof(int) * bcb.size()) != hipSuccess)
    return bail(fail(SAC_E_NOMEM, "bc alloc"));
  hipMemcpy(d->bc_e, bce.data(), sizeof(int) * bce.size(), hipMemcpyHostToDevice);
  hipMemcpy(d->bc_b, bcb.data(), sizeof(int) * bcb.size(), hipMemcpyHostToDevice);
  if (hipMalloc((void**)&d->red, sizeof(unsigned long long) * (kSlots + 1) * sac::kRedSlots) !=
      hipSuccess)
    return bail(fail(SAC_E_NOMEM, "reduction buffer"));
  hipMemset(d->red, 0, sizeof(unsigned long long) * (kSlots + 1) * sac::kRedSlots);
  // getdt_grav (src/vacusr.gravity.t:114-121): computed once, at it==itmin
  d->dtgrav = kBigDouble;
  for (int k = 0; k < ndim; k++) {
    double gk = std::fabs(d->grav[k]);
    if (gk > 0.0) {
      double mx = -kBigDouble;
      for (int j = 0; j < ndim; j++) mx = std::fmax(mx, gk / d->dxmin[j]);
      d->dtgrav = std::fmin(d->dtgrav, 1.0 / std::sqrt(mx));
    }
  }
  if (hipDeviceSynchronize() != hipSuccess) return bail(fail(SAC_E_HIP, "init sync"));
  *out = d;
  return 0;
}

int sac_dev_free(sac_handle h) {
  if (!h) return 0;
  Dev* d = (Dev*)h;
  if (d->stream) hipStreamSynchronize(d->stream);
  t_collect(*d);
  for (int s = 0; s < kSlots; s++) hipFree(d->reg[s]);
  hipFree(d->bg);
  hipFree(d->rate);
  hipFree(d->cross);
  hipFree(d->J);
  hipFree(d->snap);
  for (int k = 0; k < 3; k++) hipFree(d->dx[k]);
  hipFree(d->bc_e);
  hipFree(d->bc_b);
  hipFree(d->red);
  if (d->cl) {
    Closure* c = d->cl;
    hipFree(c->P32);
    hipFree(c->P16);
    hipFree(c->m1);
    hipFree(c->m2);
    hipFree(c->partial);
    hipFree(c->loss);
    for (int k = 0; k < 3; k++) hipFree(c->RF[k]);
    hipFree(c->feat);
    hipFree(c->targ);
    delete c;
  }
  if (d->save_stream) hipStreamSynchronize(d->save_stream);
  hipFree(d->save_dev);
  if (d->save_host) hipHostFree(d->save_host);
  if (d->save_taken) hipEventDestroy(d->save_taken);
  if (d->save_done) hipEventDestroy(d->save_done);
  if (d->save_stream) hipStreamDestroy(d->save_stream);
  if (d->stream) hipStreamDestroy(d->stream);
  delete d;
  return 0;
}

int sac_dev_upload(sac_handle h, const double* w) {
  if (int rc = check_handle(h)) return rc;
  if (!w) return fail(SAC_E_ARG, "w is null");
  Dev& d = *(Dev*)h;
  if (int rc = copy_in(d, d.reg[SAC_SLOT_W], w, d.ne, d.esz)) return rc;
  if (int rc = copy_in(d, d.bg, w + (size_t)d.ne * d.ncell_grid, d.nbg, sizeof(double))) return rc;
  d.red_valid[SAC_SLOT_W] = false;
  d.bg_host.clear();  // the ghosts of the background change at the next getboundary
  d.bg_ghosts_stale = d.initialized;
  return 0;
}

int sac_dev_download(sac_handle h, double* w) {
  if (int rc = check_handle(h)) return rc;
  if (!w) return fail(SAC_E_ARG, "w is null");
  Dev& d = *(Dev*)h;
  if (int rc = copy_out(d, w, d.reg[SAC_SLOT_W], d.ne, d.esz)) return rc;
  return copy_out(d, w + (size_t)d.ne * d.ncell_grid, d.bg, d.nbg, sizeof(double));
}

int sac_dev_upload_slot(sac_handle h, int slot, const double* w) {
  if (int rc = check_handle(h)) return rc;
  if (slot < 0 || slot >= kSlots || !w) return fail(SAC_E_ARG, "bad slot %d", slot);
  Dev& d = *(Dev*)h;
  d.red_valid[slot] = false;
  return copy_in(d, d.reg[slot], w, d.ne, d.esz);
}

int sac_dev_download_slot(sac_handle h, int slot, double* w) {
  if (int rc = check_handle(h)) return rc;
  if (slot < 0 || slot >= kSlots || !w) return fail(SAC_E_ARG, "bad slot %d", slot);
  Dev& d = *(Dev*)h;
  if (int rc = copy_out(d, w, d.reg[slot], d.ne, d.esz)) return rc;
  return copy_out(d, w + (size_t)d.ne * d.ncell_grid, d.bg, d.nbg, sizeof(double));
}

// ---------------------------------------------------------------------
// asynchronous saves: the snapshot is a device-to-device copy of W on the
// handle's stream (so later steps may overwrite W at once), copied to pinned
// host memory on the save stream after an event; sac_dev_save_end waits for
// that copy and writes the caller's array.

int sac_dev_save_begin(sac_handle h) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  if (d.save_pending) return fail(SAC_E_ARG, "a save is already in progress");
  const size_t fbytes = d.esz * (size_t)d.g.fs;
  if (!d.save_stream) {
    HIPCHK(hipStreamCreateWithFlags(&d.save_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&d.save_taken, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&d.save_done, hipEventDisableTiming));
    if (int rc = balloc(d, &d.save_dev, fbytes * (size_t)d.ne)) return rc;
    HIPCHK(hipHostMalloc(&d.save_host, d.esz * (size_t)d.ncell_grid * (size_t)d.ne,
                         hipHostMallocDefault));
  }
  if (d.bg_host.empty()) {  // the background is static: one synchronous copy
    d.bg_host.resize((size_t)d.nbg * (size_t)d.ncell_grid);
    if (int rc = copy_out(d, d.bg_host.data(), d.bg, d.nbg, sizeof(double))) return rc;
  }
  HIPCHK(hipMemcpyAsync(d.save_dev, d.reg[SAC_SLOT_W], fbytes * (size_t)d.ne,
                        hipMemcpyDeviceToDevice, d.stream));
  HIPCHK(hipEventRecord(d.save_taken, d.stream));
  HIPCHK(hipStreamWaitEvent(d.save_stream, d.save_taken, 0));
  const size_t w = d.esz * (size_t)d.g.G[0];
  const size_t spitch = d.esz * (size_t)d.g.st[1];
  const size_t rows = (size_t)d.g.G[1] * (size_t)d.g.G[2];
  for (int f = 0; f < d.ne; f++) {
    char* dst = (char*)d.save_host + d.esz * (size_t)f * (size_t)d.ncell_grid;
    const char* src = (const char*)d.save_dev + fbytes * (size_t)f;
    HIPCHK(hipMemcpy2DAsync(dst, w, src, spitch, w, rows, hipMemcpyDeviceToHost,
                            d.save_stream));
  }
  HIPCHK(hipEventRecord(d.save_done, d.save_stream));
  d.save_pending = true;
  return 0;
}

int sac_dev_save_ready(sac_handle h, int* ready) {
  if (int rc = check_handle(h)) return rc;
  if (!ready) return fail(SAC_E_ARG, "ready is null");
  Dev& d = *(Dev*)h;
  if (!d.save_pending) return fail(SAC_E_ARG, "no save in progress");
  const hipError_t e = hipEventQuery(d.save_done);
  if (e != hipSuccess && e != hipErrorNotReady)
    return fail(SAC_E_HIP, "hipEventQuery: %s", hipGetErrorString(e));
  *ready = (e == hipSuccess) ? 1 : 0;
  return 0;
}

int sac_dev_save_end(sac_handle h, double* w) {
  if (int rc = check_handle(h)) return rc;
  if (!w) return fail(SAC_E_ARG, "w is null");
  Dev& d = *(Dev*)h;
  if (!d.save_pending) return fail(SAC_E_ARG, "no save in progress");
  d.save_pending = false;
  HIPCHK(hipEventSynchronize(d.save_done));
  const size_t n = (size_t)d.ne * (size_t)d.ncell_grid;
  if (d.esz == sizeof(double)) {
    memcpy(w, d.save_host, sizeof(double) * n);
  } else {
    const float* s = (const float*)d.save_host;
    for (size_t i = 0; i < n; i++) w[i] = (double)s[i];
  }
  memcpy(w + n, d.bg_host.data(), sizeof(double) * d.bg_host.size());
  return 0;
}

int sac_dev_getboundary(sac_handle h) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  if (d.any_zero) {
    ghost_fill_sequential(d, SAC_SLOT_W, !d.initialized || d.bg_ghosts_stale);
    d.initialized = true;  // no stored values in this path
    d.bg_ghosts_stale = false;
    d.bg_host.clear();
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(d.stream));
    return 0;
  }
  if (!d.initialized || d.bg_ghosts_stale) {
    // background ghosts: on the first call (the background is static, so
    // the reference fills them once) and after an upload replaced them
    sac::GhostArgs a{};
    a.g = d.g;
    a.prec = sac::PREC_FP64;
    a.f = d.bg;
    a.bc = d.bc_b;
    a.nvar = d.nbg;
    a.snap = nullptr;
    a.keep_fixed = d.fullgridini ? 1 : 0;
    a.general = d.cont1_b || (d.ext1_b && !a.keep_fixed);
    sac::launch_ghost(a, d.stream);
    d.bg_ghosts_stale = false;
    d.bg_host.clear();  // background ghosts filled: the cached copy is stale
  }
  if (!d.initialized) {
    // first call: then store 'fixed' values
    ghost_fill_evolving(d, SAC_SLOT_W);
    if (d.any_fixed) {
      // fixed/fixed1 values and grad1 differences of the first fill
      // (src/vacgrid.t:473-507)
      sac::GhostArgs st{};
      st.g = d.g;
      st.prec = d.prec;
      st.f = d.reg[SAC_SLOT_W];
      st.bc = d.bc_e;
      st.nvar = d.ne;
      sac::launch_ghost_store(st, d.snap, d.stream);
    }
    d.initialized = true;
  } else {
    ghost_fill_evolving(d, SAC_SLOT_W);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return 0;
}

int sac_dev_getdt_courant(sac_handle h, double courantpar, double* dt, double* courantmax) {
  if (int rc = check_handle(h)) return rc;
  if (!dt) return fail(SAC_E_ARG, "dt is null");
  Dev& d = *(Dev*)h;
  if (!d.red_valid[SAC_SLOT_W]) cmax_reduce(d, SAC_SLOT_W);
  HIPCHK(hipGetLastError());
  double cmd[3] = {0.0, 0.0, 0.0};
  red_host(d, SAC_SLOT_W, d.ndim, d.ndim, cmd);  // cmax_k/dx_k
  double cm = 0.0;
  for (int k = 0; k < d.ndim; k++) cm = std::fmax(cm, cmd[k]);
  double v = kBigDouble;
  if (cm > kSmallDouble) v = std::fmin(v, courantpar / cm);
  *dt = v;
  if (courantmax) *courantmax = cm;
  return 0;
}

int sac_dev_getdt_special(sac_handle h, double* dt) {
  if (int rc = check_handle(h)) return rc;
  if (!dt) return fail(SAC_E_ARG, "dt is null");
  Dev& d = *(Dev*)h;
  double v = *dt;
  // getdt (src/vacphys.mhd0.t:83-103) -> getdt_res (src/vacphys.mhdres.t:5-44),
  // called before getdt_special in the main loop (src/vac.t:65-66)
  if (d.eta > 0.0) {
    double mn = kBigDouble;
    for (int k = 0; k < d.ndim; k++) mn = std::fmin(mn, d.dxmin[k]);
    v = std::fmin(v, d.dtdiffpar * mn * mn / d.eta);
  }
  if (d.visc_on) {
    unsigned long long u = 0;
    HIPCHK(hipMemcpyAsync(&u, red_visc(d), sizeof(u), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    double mv;
    memcpy(&mv, &u, sizeof(mv));
    // getdt_visc: dtdiffpar=0.25 (assigned to the common variable, which the
    // resistive limit of later steps uses); +Inf when maxviscoef==0
    d.dtdiffpar = 0.25;
    for (int k = 0; k < d.ndim; k++) {
      double mx = mv / (d.dxmin[k] * d.dxmin[k]);
      double dtd = 0.25 / mx;
      v = std::fmin(v, dtd);
    }
    sac::launch_zero_u64(red_visc(d), 1, d.stream);
  }
  v = std::fmin(v, d.dtgrav);
  *dt = v;
  return 0;
}

int sac_dev_advance(sac_handle h, double dt) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  if (!d.initialized) return fail(SAC_E_ARG, "getboundary must run before advance");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(SAC_E_ARG, "dt=%g", dt);
  const int W = SAC_SLOT_W, S1 = SAC_SLOT_S1, S2 = SAC_SLOT_S2, S3 = SAC_SLOT_S3;
  // The integrators of src/vac.t:406-475.  A stage is
  // out = base + qa*R(wct), followed by the ghost fill of out, as advect1
  // ends with getboundary (src/vac.t:544).
  const size_t nreg = (size_t)d.ne * d.g.fs;
  switch (d.typeadvance) {
    case SAC_ADV_FOURSTEP:
      // classical RK4 (:439-455): w1, w2, w3 in S1, S2, S3; the last stage
      // adds (w1 + 2*w2 + w3 - 4*w)/3 to w + dt/6*k4 (:450-454) as it forms it
      stage(d, W, W, S1, dt / 2);
      stage(d, S1, W, S2, dt / 2);
      stage(d, S2, W, S3, dt);
      stage(d, S3, W, W, dt / 6, true);
      break;
    case SAC_ADV_TWOSTEP:
      // predictor advect1(dt/2, w -> w1), full step advect1(dt, wCT=w1, w) (:424-428)
      stage(d, W, W, S1, dt / 2);
      stage(d, S1, W, W, dt);
      break;
    case SAC_ADV_THREESTEP:
      // Shu's third-order method (:430-436): w1 = w + dt R(w);
      // w2 = 3/4 w + 1/4 w1 + dt/4 R(w1); w = w/3 + 2/3 w2 + 2dt/3 R(w2)
      stage(d, W, W, S1, dt);
      sac::launch_axpy4(d.prec, (long long)nreg, d.reg[S2], d.reg[W], 0.75, d.reg[S1], 0.25,
                        nullptr, 0.0, nullptr, 0.0, d.stream);
      d.red_valid[S2] = false;
      stage(d, S1, S2, S2, dt / 4);
      sac::launch_axpy4(d.prec, (long long)nreg, d.reg[W], d.reg[W], 1.0 / 3.0, d.reg[S2],
                        2.0 / 3.0, nullptr, 0.0, nullptr, 0.0, d.stream);
      d.red_valid[W] = false;
      stage(d, S2, W, W, dt * 2.0 / 3.0);
      break;
    case SAC_ADV_STERCK:
      // (:458-466); 0.12 is a default-real (single precision) literal there
      stage(d, W, W, S1, dt * (double)0.12f);
      stage(d, S1, W, S2, dt / 4);
      stage(d, S2, W, S1, dt / 2);
      stage(d, S1, W, W, dt);
      break;
    case SAC_ADV_JAMESON:
      // (:468-475)
      stage(d, W, W, S1, dt / 4);
      stage(d, S1, W, S2, dt / 3);
      stage(d, S2, W, S1, dt / 2);
      stage(d, S1, W, W, dt);
      break;
    default: {
      // onestep: w1=w; advect1(dt, wCT=w1, w) (:407-408)
      HIPCHK(hipMemcpyAsync(d.reg[S1], d.reg[W], d.esz * nreg, hipMemcpyDeviceToDevice,
                            d.stream));
      d.red_valid[S1] = d.red_valid[W];
      if (d.red_valid[W])
        HIPCHK(hipMemcpyAsync(red_of(d, S1), red_of(d, W),
                              sizeof(unsigned long long) * sac::kRedSlots,
                              hipMemcpyDeviceToDevice, d.stream));
      stage(d, S1, W, W, dt);
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int sac_dev_set_resistive(sac_handle h, int compactres, double dtdiffpar) {
  if (int rc = check_handle(h)) return rc;
  if (!(dtdiffpar > 0.0)) return fail(SAC_E_ARG, "dtdiffpar must be > 0");
  Dev& d = *(Dev*)h;
  d.compactres = compactres ? 1 : 0;
  d.dtdiffpar = dtdiffpar;
  return 0;
}

int sac_dev_set_plane_tiles(sac_handle h, int on) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  d.plane = on ? 1 : 0;
  return 0;
}

int sac_dev_set_divbfix(sac_handle h, int on) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  if (on && d.ndim > 1 && !sac::line_tiles_built())
    return fail(SAC_E_UNSUPPORTED, "divbfix needs the line-tile kernels");
  // the reference adds nothing in 1-D ({^NOONED ...}, src/vacphys.mhd0.t:256)
  d.divbfix = (on && d.ndim > 1) ? 1 : 0;
  return 0;
}

int sac_dev_advect1(sac_handle h, double qdt, int wct_slot, int w_slot) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  if (wct_slot < 0 || wct_slot >= kSlots || w_slot < 0 || w_slot >= kSlots || wct_slot == w_slot)
    return fail(SAC_E_ARG, "advect1 needs two distinct slots (%d,%d)", wct_slot, w_slot);
  if (!d.initialized) return fail(SAC_E_ARG, "getboundary must run before advect1");
  stage(d, wct_slot, w_slot, w_slot, qdt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return 0;
}

int sac_dev_copy_slot(sac_handle h, int dst_slot, int src_slot) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  if (dst_slot < 0 || dst_slot >= kSlots || src_slot < 0 || src_slot >= kSlots)
    return fail(SAC_E_ARG, "bad slots %d %d", dst_slot, src_slot);
  if (dst_slot == src_slot) return 0;
  hipEvent_t e0{};
  t_begin(d, SAC_K_COPY, &e0);
  sac::launch_axpy4(d.prec, (long long)d.ne * d.g.fs, d.reg[dst_slot], d.reg[src_slot], 1.0,
                    nullptr, 0.0, nullptr, 0.0, nullptr, 0.0, d.stream);
  t_end(d, SAC_K_COPY, e0);
  d.red_valid[dst_slot] = false;
  HIPCHK(hipGetLastError());
  return 0;
}

int sac_dev_get_maxviscoef(sac_handle h, double* v) {
  if (int rc = check_handle(h)) return rc;
  if (!v) return fail(SAC_E_ARG, "null");
  Dev& d = *(Dev*)h;
  unsigned long long u = 0;
  HIPCHK(hipMemcpyAsync(&u, red_visc(d), sizeof(u), hipMemcpyDeviceToHost, d.stream));
  HIPCHK(hipStreamSynchronize(d.stream));
  memcpy(v, &u, sizeof(double));
  return 0;
}

int sac_dev_set_maxviscoef(sac_handle h, double v) {
  if (int rc = check_handle(h)) return rc;
  if (!(v >= 0.0)) return fail(SAC_E_ARG, "maxviscoef must be >= 0");
  Dev& d = *(Dev*)h;
  unsigned long long u;
  memcpy(&u, &v, sizeof(u));
  HIPCHK(hipMemcpyAsync(red_visc(d), &u, sizeof(u), hipMemcpyHostToDevice, d.stream));
  HIPCHK(hipStreamSynchronize(d.stream));
  return 0;
}

int sac_dev_timing(sac_handle h, int enable) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  hipStreamSynchronize(d.stream);
  t_collect(d);
  for (int k = 0; k < 10; k++) {
    d.t_ms[k] = 0.0;
    d.t_n[k] = 0;
  }
  d.timing = enable != 0;
  return 0;
}

int sac_dev_timing_get(sac_handle h, int kernel, double* total_ms, long long* count) {
  if (int rc = check_handle(h)) return rc;
  if (kernel < 0 || kernel > 9) return fail(SAC_E_ARG, "kernel id %d", kernel);
  Dev& d = *(Dev*)h;
  t_collect(d);
  if (total_ms) *total_ms = d.t_ms[kernel];
  if (count) *count = d.t_n[kernel];
  return 0;
}

int sac_dev_synchronize(sac_handle h) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  HIPCHK(hipStreamSynchronize(d.stream));
  HIPCHK(hipGetLastError());
  return 0;
}

int sac_dev_memory(sac_handle h, long long* bytes) {
  if (int rc = check_handle(h)) return rc;
  if (bytes) *bytes = ((Dev*)h)->bytes;
  return 0;
}

/* ---------------------------------------------------------------------
 * learned hyperdiffusion closure */

int sac_closure_nparam(void) { return sac::mlp::NPARAM; }

int sac_closure_init(sac_handle h, long long capacity, int max_blocks, unsigned seed) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  if (d.cl) return fail(SAC_E_ARG, "closure already initialised");
  if (d.prec != SAC_PREC_FP64) return fail(SAC_E_UNSUPPORTED, "the learned closure runs in FP64 storage");
  if (capacity < 0 || max_blocks < 1) return fail(SAC_E_ARG, "capacity/max_blocks");
  Closure* c = new Closure();
  d.cl = c;
  const int NP = sac::mlp::NPARAM;
  int rc = 0;
  if ((rc = balloc(d, (void**)&c->P32, sizeof(float) * NP)) ||
      (rc = balloc(d, (void**)&c->P16, sizeof(uint16_t) * NP)) ||
      (rc = balloc(d, (void**)&c->m1, sizeof(float) * NP)) ||
      (rc = balloc(d, (void**)&c->m2, sizeof(float) * NP)) ||
      (rc = balloc(d, (void**)&c->partial, sizeof(float) * (size_t)NP * max_blocks)) ||
      (rc = balloc(d, (void**)&c->loss, sizeof(float) * max_blocks)))
    return rc;
  c->max_blocks = max_blocks;
  const int np_var = 2 * d.ndim + 1;
  for (int k = 0; k < d.ndim; k++)
    if ((rc = balloc(d, (void**)&c->RF[k], sizeof(float) * (size_t)np_var * d.g.fs))) return rc;
  if (capacity > 0) {
    if ((rc = balloc(d, (void**)&c->feat, sizeof(uint16_t) * sac::closure::kFeat * (size_t)capacity)) ||
        (rc = balloc(d, (void**)&c->targ, sizeof(float) * sac::closure::kTarg * (size_t)capacity)))
      return rc;
  }
  c->cap = capacity;
  // He-normal initial weights (fan-in), zero biases, from a seeded
  // xorshift + Box-Muller stream
  std::vector<float> p(NP, 0.f);
  unsigned long long st = 0x9E3779B97F4A7C15ull ^ (unsigned long long)seed;
  auto uni = [&]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int rows, int cols) {
    const double sd = std::sqrt(2.0 / cols);
    for (int i = 0; i < rows * cols; i++) {
      const double u1 = uni(), u2 = uni();
      p[off + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
  };
  using namespace sac::mlp;
  fill(OFF_W1, H, IN);
  fill(OFF_W2, H, H);
  fill(OFF_W3, H, H);
  fill(OFF_W4, OUT, H);
  HIPCHK(hipMemcpy(c->P32, p.data(), sizeof(float) * NP, hipMemcpyHostToDevice));
  sac::closure::launch_to_bf16(c->P32, c->P16, d.stream);
  HIPCHK(hipStreamSynchronize(d.stream));
  return 0;
}

static int closure_of(sac_handle h, Closure** out) {
  if (int rc = check_handle(h)) return rc;
  Dev& d = *(Dev*)h;
  if (!d.cl) return fail(SAC_E_ARG, "sac_closure_init has not run on this handle");
  *out = d.cl;
  return 0;
}

int sac_closure_set_params(sac_handle h, const float* p, int n) {
  Closure* c;
  if (int rc = closure_of(h, &c)) return rc;
  if (!p || n != sac::mlp::NPARAM) return fail(SAC_E_ARG, "need %d parameters", sac::mlp::NPARAM);
  Dev& d = *(Dev*)h;
  HIPCHK(hipMemcpy(c->P32, p, sizeof(float) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->m1, 0, sizeof(float) * n));
  HIPCHK(hipMemset(c->m2, 0, sizeof(float) * n));
  c->adam_t = 0;
  sac::closure::launch_to_bf16(c->P32, c->P16, d.stream);
  HIPCHK(hipStreamSynchronize(d.stream));
  return 0;
}

int sac_closure_get_params(sac_handle h, float* p, int n) {
  Closure* c;
  if (int rc = closure_of(h, &c)) return rc;
  if (!p || n != sac::mlp::NPARAM) return fail(SAC_E_ARG, "need %d parameters", sac::mlp::NPARAM);
  HIPCHK(hipMemcpy(p, c->P32, sizeof(float) * n, hipMemcpyDeviceToHost));
  return 0;
}

int sac_closure_collect(sac_handle h, int slot, long long* total) {
  Closure* c;
  if (int rc = closure_of(h, &c)) return rc;
  Dev& d = *(Dev*)h;
  if (slot < 0 || slot >= kSlots) return fail(SAC_E_ARG, "bad slot %d", slot);
  long long nm = 1;
  for (int k = 0; k < 3; k++) nm *= d.g.Mhi[k] - d.g.Mlo[k] + 1;
  const long long add = nm * d.ndim;
  if (c->n + add > c->cap)
    return fail(SAC_E_ARG, "sample buffer full (%lld + %lld > %lld)", c->n, add, c->cap);
  sac::closure::launch_collect(d.ndim, d.ndir, d.g, (const double*)d.reg[slot], d.bg, d.gamma,
                               c->sensor, c->feat, c->targ, c->n, d.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  c->n += add;
  if (total) *total = c->n;
  return 0;
}

int sac_closure_collect_from(sac_handle h, sac_handle src, int slot, long long* total) {
  Closure* c;
  if (int rc = closure_of(h, &c)) return rc;
  if (int rc = check_handle(src)) return rc;
  Dev& d = *(Dev*)h;
  Dev& s = *(Dev*)src;
  if (slot < 0 || slot >= kSlots) return fail(SAC_E_ARG, "bad slot %d", slot);
  if (s.prec != SAC_PREC_FP64) return fail(SAC_E_UNSUPPORTED, "samples come from FP64 storage");
  if (s.device != d.device) return fail(SAC_E_ARG, "the two handles are on different devices");
  long long nm = 1;
  for (int k = 0; k < 3; k++) nm *= s.g.Mhi[k] - s.g.Mlo[k] + 1;
  const long long add = nm * s.ndim;
  if (c->n + add > c->cap)
    return fail(SAC_E_ARG, "sample buffer full (%lld + %lld > %lld)", c->n, add, c->cap);
  HIPCHK(hipStreamSynchronize(s.stream));  // the source state is complete
  sac::closure::launch_collect(s.ndim, s.ndir, s.g, (const double*)s.reg[slot], s.bg, s.gamma,
                               c->sensor, c->feat, c->targ, c->n, d.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  c->n += add;
  if (total) *total = c->n;
  return 0;
}

int sac_closure_clear(sac_handle h) {
  Closure* c;
  if (int rc = closure_of(h, &c)) return rc;
  c->n = 0;
  return 0;
}

int sac_closure_train(sac_handle h, int steps, int blocks, double lr, long long base,
                      long long stride, double* mean_loss) {
  Closure* c;
  if (int rc = closure_of(h, &c)) return rc;
  Dev& d = *(Dev*)h;
  if (c->n < 64) return fail(SAC_E_ARG, "no training samples");
  if (blocks < 1 || blocks > c->max_blocks) return fail(SAC_E_ARG, "blocks %d", blocks);
  if (stride < 1) return fail(SAC_E_ARG, "stride");
  const long long per = (long long)blocks * sac::mlp::WAVE;
  double lsum = 0.0;
  std::vector<float> lb(blocks);
  for (int s = 0; s < steps; s++) {
    sac::closure::launch_train(c->feat, c->targ, c->n, base + (long long)s * per, stride, c->P16,
                               c->P32, c->partial, c->loss, blocks, 1.0f / (float)per, d.stream);
    c->adam_t += 1;
    sac::closure::launch_adam(c->P32, c->P16, c->m1, c->m2, c->partial, blocks, (float)lr, 0.9f,
                              0.999f, 1e-8f, c->adam_t, d.stream);
    if (mean_loss && s == steps - 1) {
      HIPCHK(hipMemcpyAsync(lb.data(), c->loss, sizeof(float) * blocks, hipMemcpyDeviceToHost,
                            d.stream));
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  if (mean_loss) {
    for (int b = 0; b < blocks; b++) lsum += lb[b];
    *mean_loss = lsum / (double)per;
  }
  return 0;
}

int sac_closure_grad(sac_handle h, const float* feat, const float* targ, int nsamples,
                     float* grad, double* loss) {
  Closure* c;
  if (int rc = closure_of(h, &c)) return rc;
  Dev& d = *(Dev*)h;
  if (!feat || !targ || !grad || nsamples < 64 || nsamples % 64 != 0)
    return fail(SAC_E_ARG, "nsamples must be a positive multiple of 64");
  const int blocks = nsamples / 64;
  if (blocks > c->max_blocks) return fail(SAC_E_ARG, "too many samples");
  const int NP = sac::mlp::NPARAM;
  uint16_t* f16 = nullptr;
  float* tg = nullptr;
  HIPCHK(hipMalloc((void**)&f16, sizeof(uint16_t) * sac::closure::kFeat * nsamples));
  HIPCHK(hipMalloc((void**)&tg, sizeof(float) * sac::closure::kTarg * nsamples));
  // features to BF16 with round-to-nearest-even, as the sample kernel does
  std::vector<uint16_t> hb((size_t)sac::closure::kFeat * nsamples);
  for (size_t i = 0; i < hb.size(); i++) {
    uint32_t u;
    memcpy(&u, &feat[i], 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    hb[i] = (uint16_t)(u >> 16);
  }
  int rc = 0;
  if (hipMemcpy(f16, hb.data(), sizeof(uint16_t) * hb.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(tg, targ, sizeof(float) * sac::closure::kTarg * nsamples, hipMemcpyHostToDevice) !=
          hipSuccess)
    rc = fail(SAC_E_HIP, "upload");
  if (!rc) {
    sac::closure::launch_train(f16, tg, nsamples, 0, 1, c->P16, c->P32, c->partial, c->loss,
                               blocks, 1.0f / (float)nsamples, d.stream);
    std::vector<float> part((size_t)NP * blocks), lb(blocks);
    if (hipStreamSynchronize(d.stream) != hipSuccess ||
        hipMemcpy(part.data(), c->partial, sizeof(float) * part.size(), hipMemcpyDeviceToHost) !=
            hipSuccess ||
        hipMemcpy(lb.data(), c->loss, sizeof(float) * blocks, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(SAC_E_HIP, "training kernel: %s", hipGetErrorString(hipGetLastError()));
    } else {
      for (int k = 0; k < NP; k++) {
        double g = 0.0;
        for (int b = 0; b < blocks; b++) g += part[(size_t)b * NP + k];
        grad[k] = (float)g;
      }
      double ls = 0.0;
      for (int b = 0; b < blocks; b++) ls += lb[b];
      if (loss) *loss = ls / nsamples;
    }
  }
  hipFree(f16);
  hipFree(tg);
  return rc;
}

int sac_closure_enable(sac_handle h, int on, double sensor) {
  Closure* c;
  if (int rc = closure_of(h, &c)) return rc;
  if (!(sensor > 0.0)) return fail(SAC_E_ARG, "sensor threshold must be > 0");
  if (on < 0 || on > 2) return fail(SAC_E_ARG, "on must be 0, 1 or 2");
  c->on = on;
  c->sensor = (float)sensor;
  return 0;
}

}  // extern "C"
