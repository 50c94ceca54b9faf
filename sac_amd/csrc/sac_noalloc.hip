// sac_noalloc.hip -- getdt_courant on a state the caller keeps in device
// memory, without any device allocation (for callers that cannot allocate:
// code under HIP graph capture, serving loops, harnesses that forbid it).
//
// It runs the kernels of sac_dev_getdt_courant (the cmax max-reduction of
// getdt_courant, src/vac.t:641-699, with getcmax, src/vacphys0.t.mhd:77-112)
// on the caller's arrays in the reference layout; the reduction slots that
// sac_dev_getdt_courant keeps in its handle's buffer live here in a
// file-scope device array instead.  Every call on a device shares that
// device's array, so calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/sac_amd.h"
#include "sac_kernels.h"

namespace sac {
int set_error(int code, const char* fmt, ...);
}

namespace {

constexpr int kMaxDevices = 64;
// reduction slots: [d] = max cmax_d, [ndim+d] = max cmax_d/dx_d (d < ndim)
constexpr int kScratchSlots = 2 * sac::kMaxDim;
// largest grid (cells including ghosts) a call accepts: the cmax launch
// numbers its threads with 32-bit block indices
constexpr long long kMaxGridCells = 1LL << 36;

__device__ __align__(256) unsigned long long scratch[kScratchSlots];

std::mutex g_mu;
unsigned long long* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// address of this device's copy of `scratch` (looked up once per device,
// with that device current)
int scratch_of(int dev, unsigned long long** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return sac::set_error(SAC_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[dev] = (unsigned long long*)p;
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" int sac_getdt_courant_noalloc(int ndim, int ndir, const int* nx, int precision,
                                         const void* w, const double* bg, const double* dx,
                                         double gamma, double courantpar, void* stream,
                                         double* dt, double* courantmax) {
  if (ndim < 1 || ndim > 3 || ndir < ndim || ndir > 3)
    return sac::set_error(SAC_E_ARG, "need 1<=ndim<=ndir<=3 (ndim=%d ndir=%d)", ndim, ndir);
  if (!nx || !w || !bg || !dx || !dt) return sac::set_error(SAC_E_ARG, "null argument");
  if (precision != SAC_PREC_FP64 && precision != SAC_PREC_MIXED)
    return sac::set_error(SAC_E_UNSUPPORTED, "precision %d not built", precision);
  if (!(gamma > 0.0)) return sac::set_error(SAC_E_ARG, "gamma must be > 0");
  sac::GridDesc g{};
  g.ndim = ndim;
  g.ndir = ndir;
  long long cells = 1;
  for (int k = 0; k < 3; k++) {
    if (k < ndim) {
      if (nx[k] < 6) return sac::set_error(SAC_E_ARG, "nx[%d]=%d < 6", k, nx[k]);
      g.G[k] = nx[k] + 4;
      g.Mlo[k] = 2;
      g.Mhi[k] = nx[k] + 1;
    } else {
      g.G[k] = 1;
      g.Mlo[k] = 0;
      g.Mhi[k] = 0;
    }
    cells *= g.G[k];
    if (cells > kMaxGridCells)
      return sac::set_error(SAC_E_ARG, "grid of more than %lld cells", kMaxGridCells);
  }
  // the reference layout: unpadded rows, fields one after another
  g.st[0] = 1;
  g.st[1] = g.G[0];
  g.st[2] = (long long)g.G[0] * g.G[1];
  g.fs = cells;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return sac::set_error(SAC_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return sac::set_error(SAC_E_UNSUPPORTED, "device %d beyond the %d supported", dev, kMaxDevices);
  unsigned long long* red = nullptr;
  if (int rc = scratch_of(dev, &red)) return rc;
  hipStream_t st = (hipStream_t)stream;
  sac::launch_zero_u64(red, kScratchSlots, st);  // a fill kernel (no memset node)
  sac::CmaxArgs a{};
  a.g = g;
  a.prec = precision;
  a.w = w;
  a.bg = bg;
  a.dx[0] = dx;
  a.dx[1] = (ndim > 1) ? dx + g.G[0] : nullptr;
  a.dx[2] = (ndim > 2) ? dx + g.G[0] + g.G[1] : nullptr;
  a.gamma = gamma;
  a.red_out = red;
  sac::launch_cmax(a, st);
  e = hipGetLastError();
  if (e != hipSuccess) return sac::set_error(SAC_E_HIP, "cmax launch: %s", hipGetErrorString(e));
  unsigned long long v[kScratchSlots] = {};
  e = hipMemcpyAsync(v, red, sizeof(v), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return sac::set_error(SAC_E_HIP, "reduction copy: %s", hipGetErrorString(e));
  // the arithmetic of sac_dev_getdt_courant
  double cm = 0.0;
  for (int k = 0; k < ndim; k++) {
    double x;
    memcpy(&x, &v[ndim + k], sizeof(x));
    cm = std::fmax(cm, x);
  }
  double r = 1.0e99;
  if (cm > 1.0e-99) r = std::fmin(r, courantpar / cm);
  *dt = r;
  if (courantmax) *courantmax = cm;
  return 0;
}
