// sac_batch.hip -- ensembles of independent 1-D MHD tubes of uneven length,
// or of 2-D MHD patches of uneven size, advanced together on one GPU.
//
// Every tube (or patch) is its own serial run of the reference
// (src/vac.t:55-106): its own dt from getdt_courant and getdt_special
// (lagged maxviscoef, forced dtdiffpar 0.25), its own t and it, 'cont'
// boundaries, classical RK4 ('fourstep').  The tubes share the kernels: one
// launch per kernel covers every tube through a tile table (tube, first
// mesh cell along x and y), tiles never straddle two tubes, so the per-tube
// reductions are block reductions plus one atomic per block.  The time loop
// runs on the device: a small kernel computes each tube's next dt and
// finish flag, the RK stages use it, and a second small kernel advances t
// and it.  One step is a fixed sequence of launches, captured once in a
// hipGraph holding several steps and replayed until every tube has reached
// tmax; finished tubes skip their tiles.
//
// The line work and the combining update are the grid path's own device
// functions (line1_tile for tubes, line_tile for patches, final_cell)
// applied to a per-tube grid descriptor, so a tube or patch advanced here
// follows the same arithmetic as the same tube or patch run alone through
// sac_dev_*, bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sac_amd.h"
#include "sac_device.h"
#include "sac_line.h"

namespace sac {
namespace batch {

constexpr int NDIR = 3;
constexpr int NE = 2 + 2 * NDIR;  // evolving fields
constexpr int NBG = 2 + NDIR;     // background fields
constexpr int NT2 = 4;            // cross-term factor fields of a 2-D patch (2*ndim*(ndim-1))
constexpr int kSlots = 4;         // W, S1, S2, S3
constexpr int kRed = 4;           // per tube and register: max cmax_d, then max cmax_d/dx_d (d < ndim)

// FINISHED: t >= tmax or it >= itmax; STOPPED: dt < dtmin (the reference's
// third finish condition, src/vac.t:591, hit when a tube's state blows up
// and cmax or maxviscoef become infinite); FAILED: dt not a number
enum TubeState : int { RUNNING = 0, FINISHED = 1, FAILED = 2, STOPPED = 3 };

struct Tile {
  int tube;  // tube index
  int c0;    // first mesh cell (grid index) of the tile along x
  int c1;    // and along y (2-D patches; 0 for tubes)
};

struct Args {
  int ndim;                  // 1: tubes, 2: patches
  int ntubes;
  long long fs;              // field stride (elements)
  const long long* start;    // grid origin of each tube in the packed layout
  const int* nxy;            // mesh cells per tube along x and y: [tube][2]
  const int* pitch;          // row pitch (elements) of each tube's grid
  const long long* dxo;      // offset of each tube's cell widths in dx (x, then y)
  const Tile* tiles;         // tiles of the tubes in `active`
  int ntiles;
  int tbx;                   // tile width along x in cells (2-D)
  const int* active;         // tubes still running when the tiles were built
  int nactive;
  const Tile* ltiles[2];     // patches: line tiles of x and y as (patch, logical block, segment)
  int nltiles[2];
  double* reg[kSlots];
  const double* bg;
  const double* dx;          // cell widths per tube: G_x values, then G_y values
  double* R;
  double* T;                 // cross-term factors (2-D), packed like NT2 fields
  unsigned long long* red;   // [tube][slot][kRed]
  unsigned long long* visc;  // [tube] maxviscoef since the last dt
  double* t;
  double* dt;
  int* it;
  int* state;
  const double* dxmin;       // per tube and axis, over its mesh: [tube][2]
  double gamma;
  int visc_on;
  double courantpar, tmax, dtmin;
  int itmax;
  unsigned int* running;     // count of tubes still running after a step
};

template <int NDIM>
__device__ __forceinline__ GridDesc tube_grid(const Args& a, int tube) {
  GridDesc g{};
  g.ndim = NDIM;
  g.ndir = NDIR;
  const int nx = a.nxy[2 * tube];
  g.G[0] = nx + 4;
  g.G[1] = (NDIM == 2) ? a.nxy[2 * tube + 1] + 4 : 1;
  g.G[2] = 1;
  g.Mlo[0] = 2;
  g.Mhi[0] = nx + 1;
  if (NDIM == 2) {
    g.Mlo[1] = 2;
    g.Mhi[1] = g.G[1] - 3;
  }
  g.st[0] = 1;
  g.st[1] = a.pitch[tube];
  g.st[2] = g.st[1] * g.G[1];
  g.fs = a.fs;
  return g;
}

__device__ __forceinline__ unsigned long long* red_of(const Args& a, int tube, int slot) {
  return a.red + ((size_t)tube * kSlots + slot) * kRed;
}

// cell of thread `tid`: cell j of its group of CK along x (tubes), or its
// position in the tile (patches, CK = 1); false outside the mesh
template <int NDIM, int CK>
__device__ __forceinline__ bool tile_cell(const Args& a, const Tile& tl, const GridDesc& g, int j,
                                          int (&c)[3]) {
  const int tid = (int)threadIdx.x;
  if (NDIM == 1) {
    c[0] = tl.c0 + tid * CK + j;
    c[1] = c[2] = 0;
    return c[0] <= g.Mhi[0];
  }
  c[0] = tl.c0 + tid % a.tbx;
  c[1] = tl.c1 + tid / a.tbx;
  c[2] = 0;
  return c[0] <= g.Mhi[0] && c[1] <= g.Mhi[1];
}

// Patches: the work of axis K as the grid path's LDS line tiles (line_tile,
// sac_line.h) on the patch's own grid; ltiles[K] holds (patch, logical
// block, segment length along K).  The y kernel adds the y derivatives of
// the x kernel's cross factors, as on a grid.  The x kernel's first block
// of a patch also clears the patch's reductions of register `out`.
template <int K>
__global__ void __launch_bounds__((line::Tile<K, double>::T)) kb_pline(Args a, int wct, int out) {
  __shared__ double L[line::lds_elems<2, NDIR, K, double, false>()];
  const Tile tl = a.ltiles[K][blockIdx.x];
  if (a.state[tl.tube] != RUNNING) return;  // uniform over the block
  if (K == 0 && threadIdx.x < kRed && tl.c0 == 0) red_of(a, tl.tube, out)[threadIdx.x] = 0ull;
  const GridDesc g = tube_grid<2>(a, tl.tube);
  const long long o = a.start[tl.tube];
  AxisArgs x{};
  x.g = g;
  x.prec = PREC_FP64;
  x.wct = a.reg[wct] + o;
  x.bg = a.bg + o;
  x.dx[0] = a.dx + a.dxo[tl.tube];
  x.dx[1] = x.dx[0] + g.G[0];
  x.R = a.R + o;
  x.T = a.T + o;
  x.gamma = a.gamma;
  x.visc_on = a.visc_on;
  x.ctot_src = red_of(a, tl.tube, wct);
  x.red_visc = a.visc + tl.tube;
  x.xcd = 1;
  x.cross_last = (K == 1) ? 1 : 0;
  double numax = 0.0;
  line::line_tile<2, NDIR, K, double, double>(x, tl.c1, L, numax, tl.c0);
  if (a.visc_on) {
    double r[1] = {numax};
    block_max_atomic<1>(r, x.red_visc);
  }
}

// 'cont' ghost cells (getboundary, src/vacgrid.t:366-370) whose value is
// that of mesh cell c of register f, written by the thread that stored c.
// The reference fills the x regions over the whole grid along y, then the
// y regions, so every ghost cell takes the mesh cell at its clamped
// indices: c supplies the ghosts beyond each edge it lies on, including
// the corners.
template <int NDIM>
__device__ __forceinline__ void cont_ghosts_of(const GridDesc& g, double* f, const int (&c)[3]) {
  int gx[3] = {c[0], 0, 0}, gy[3] = {c[1], 0, 0};
  int nxg = 1, nyg = 1;
  if (c[0] == g.Mlo[0]) gx[nxg++] = 0, gx[nxg++] = 1;
  if (c[0] == g.Mhi[0]) gx[nxg++] = g.Mhi[0] + 1, gx[nxg++] = g.Mhi[0] + 2;
  if (NDIM == 2) {
    if (c[1] == g.Mlo[1]) gy[nyg++] = 0, gy[nyg++] = 1;
    if (c[1] == g.Mhi[1]) gy[nyg++] = g.Mhi[1] + 1, gy[nyg++] = g.Mhi[1] + 2;
  }
  if (nxg == 1 && nyg == 1) return;
  const long long ic = lin(g, c[0], c[1], 0);
  for (int n = 0; n < NE; n++) {
    const double v = f[n * g.fs + ic];
    for (int j = 0; j < nyg; j++)
      for (int i = 0; i < nxg; i++)
        if (i || j) f[n * g.fs + lin(g, gx[i], gy[j], 0)] = v;
  }
}

// The earlier stage registers of RK4's last stage (RK4L), as in the grid
// path: out = (base + dt/da*R) + (S1 + 2*S2 + S3 - 4*base)/3, with S3 the
// stage's input wct
template <bool RK4L>
__device__ __forceinline__ void set_rk4(const Args& a, long long o, FinalArgs& f) {
  if constexpr (RK4L) {
    f.w1 = a.reg[SAC_SLOT_S1] + o;
    f.w2 = a.reg[SAC_SLOT_S2] + o;
  }
}

// Patches: combining update of one tile: out = base + dt/da * R (see
// set_rk4), with the x derivatives of the cross factors, then the 'cont'
// ghosts of out
template <int NDIM, int CK, bool RK4L>
__global__ void __launch_bounds__(256) kb_final(Args a, int wct, int base, int out, double da) {
  const Tile tl = a.tiles[blockIdx.x];
  if (a.state[tl.tube] != RUNNING) return;
  const GridDesc g = tube_grid<NDIM>(a, tl.tube);
  const long long o = a.start[tl.tube];
  const double dt = a.dt[tl.tube];
  FinalArgs f{};
  f.g = g;
  f.prec = PREC_FP64;
  f.wct = a.reg[wct] + o;
  f.bg = a.bg + o;
  f.dx[0] = a.dx + a.dxo[tl.tube];
  if (NDIM == 2) f.dx[1] = f.dx[0] + g.G[0];
  f.R = a.R + o;
  f.T = (NDIM == 2) ? a.T + o : nullptr;
  f.base = a.reg[base] + o;
  f.out = a.reg[out] + o;
  set_rk4<RK4L>(a, o, f);
  f.qa = dt / da;
  f.gamma = a.gamma;
  f.visc_on = a.visc_on;
  f.skip_last = (NDIM == 2) ? 1 : 0;  // the y line tiles added their cross terms
  double redv[2 * NDIM];
#pragma unroll
  for (int k = 0; k < 2 * NDIM; k++) redv[k] = 0.0;
#pragma unroll
  for (int j = 0; j < CK; j++) {
    int c[3];
    if (tile_cell<NDIM, CK>(a, tl, g, j, c)) {
      double r[2 * NDIM];
      final_cell<NDIM, NDIR, double, double, RK4L>(f, g, c, r);
      cont_ghosts_of<NDIM>(g, (double*)f.out, c);
#pragma unroll
      for (int k = 0; k < 2 * NDIM; k++) redv[k] = fmax(redv[k], r[k]);
    }
  }
  block_max_atomic<2 * NDIM>(redv, red_of(a, tl.tube, out));
}

// Tubes: one whole RK stage of a tile of LT cells in one kernel: the x work
// as an LDS line tile (line1_tile, sac_device.h, also the x kernel of 1-D
// FP64 grids), then the combining update of the tile's cells (final_cell;
// the thread that stored R of its cell reads it back, so no other block is
// waited for) with the 'cont' ghosts.  The reductions of `out` must have
// been cleared by kb_zero_red.
template <int LT, bool RK4L>
__global__ void __launch_bounds__(LT) kb_tube_stage(Args a, int wct, int base, int out, double da) {
  __shared__ double L[line1::LV<NDIR> * (LT + 6)];
  const Tile tl = a.tiles[blockIdx.x];
  if (a.state[tl.tube] != RUNNING) return;  // uniform over the block
  const GridDesc g = tube_grid<1>(a, tl.tube);
  const long long o = a.start[tl.tube];
  double numax = 0.0;
  line1_tile<LT, NDIR>(g, a.reg[wct] + o, a.bg + o, a.dx + a.dxo[tl.tube], a.R + o, a.gamma,
                       a.visc_on, red_get(red_of(a, tl.tube, wct), 0), tl.c0, L, numax);
  __syncthreads();  // orders each thread's R stores before its loads below
  const double dt = a.dt[tl.tube];
  FinalArgs f{};
  f.g = g;
  f.prec = PREC_FP64;
  f.wct = a.reg[wct] + o;
  f.bg = a.bg + o;
  f.dx[0] = a.dx + a.dxo[tl.tube];
  f.R = a.R + o;
  f.base = a.reg[base] + o;
  f.out = a.reg[out] + o;
  set_rk4<RK4L>(a, o, f);
  f.qa = dt / da;
  f.gamma = a.gamma;
  f.visc_on = a.visc_on;
  double redv[2] = {0.0, 0.0};
  int c[3] = {tl.c0 + (int)threadIdx.x, 0, 0};
  if (c[0] <= g.Mhi[0]) {
    final_cell<1, NDIR, double, double, RK4L>(f, g, c, redv);
    cont_ghosts_of<1>(g, (double*)f.out, c);
  }
  if (a.visc_on) {
    double r[1] = {numax};
    block_max_atomic<1>(r, a.visc + tl.tube);
  }
  block_max_atomic<2>(redv, red_of(a, tl.tube, out));
}

// cmax reductions of register `slot` of every running tube
template <int NDIM, int CK>
__global__ void __launch_bounds__(256) kb_cmax(Args a, int slot) {
  const Tile tl = a.tiles[blockIdx.x];
  if (a.state[tl.tube] != RUNNING) return;
  const GridDesc g = tube_grid<NDIM>(a, tl.tube);
  const long long o = a.start[tl.tube];
  const double* dx0 = a.dx + a.dxo[tl.tube];
  double redv[2 * NDIM];
#pragma unroll
  for (int k = 0; k < 2 * NDIM; k++) redv[k] = 0.0;
#pragma unroll
  for (int j = 0; j < CK; j++) {
    int c[3];
    if (tile_cell<NDIM, CK>(a, tl, g, j, c)) {
      Pt<NDIR, double> p;
      load_pt<NDIR, double, double>(a.reg[slot] + o, a.bg + o, a.fs, lin(g, c[0], c[1], c[2]), p);
      double cm[NDIM];
      cmax_all<NDIM, NDIR>(p, a.gamma, cm);
#pragma unroll
      for (int d = 0; d < NDIM; d++) {
        const double dxd = (d == 0) ? dx0[c[0]] : dx0[g.G[0] + c[1]];
        redv[d] = fmax(redv[d], cm[d]);
        redv[NDIM + d] = fmax(redv[NDIM + d], cm[d] / dxd);
      }
    }
  }
  block_max_atomic<2 * NDIM>(redv, red_of(a, tl.tube, slot));
}

// 'cont' ghost cells of nfield packed fields f of the tubes: 2 per side,
// copies of the edge cell (getboundary, src/vacgrid.t:366-370), for the
// initial fill
__global__ void kb_ghost(Args a, double* f, int nfield) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = (int)(i & 3);
  if ((i >> 2) >= a.nactive) return;
  const int tube = a.active[i >> 2];
  if (a.state[tube] != RUNNING) return;
  const int n = a.nxy[2 * tube];
  const long long o = a.start[tube];
  const int cg = (k < 2) ? k : n + k;  // ghost cells 0,1 and n+2,n+3
  const int ce = (k < 2) ? 2 : n + 1;  // edge mesh cell
  for (int v = 0; v < nfield; v++) f[v * a.fs + o + cg] = f[v * a.fs + o + ce];
}

// 'cont' ghost cells of a 2-D patch (blockIdx.y = entry of the active
// list), for the initial fill.  The reference fills the x regions over the whole grid along y, then the y regions over
// the whole grid along x (src/vacgrid.t:195-512), so every ghost cell ends
// up with the mesh cell at its clamped indices.  The ghost frame of a
// G0 x G1 grid: rows 0, 1, G1-2, G1-1 in full (4*G0 cells), then columns
// 0, 1, G0-2, G0-1 of the mesh rows (4*(G1-4) cells).
__global__ void kb_ghost2(Args a, double* f, int nfield) {
  const int tube = a.active[blockIdx.y];
  if (a.state[tube] != RUNNING) return;
  const int nx = a.nxy[2 * tube], ny = a.nxy[2 * tube + 1];
  const int G0 = nx + 4, G1 = ny + 4;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= 4 * G0 + 4 * ny) return;
  int c0, c1;
  if (i < 4 * G0) {
    const int r = i / G0;
    c0 = i - r * G0;
    c1 = (r < 2) ? r : G1 - 4 + r;
  } else {
    const int j = i - 4 * G0, k = j & 3;
    c1 = 2 + (j >> 2);
    c0 = (k < 2) ? k : G0 - 4 + k;
  }
  const int e0 = min(max(c0, 2), nx + 1), e1 = min(max(c1, 2), ny + 1);
  const long long p = a.pitch[tube];
  const long long dst = a.start[tube] + c0 + c1 * p, src = a.start[tube] + e0 + e1 * p;
  for (int v = 0; v < nfield; v++) f[v * a.fs + dst] = f[v * a.fs + src];
}

// zero the reductions of register `slot` of every tube
__global__ void kb_zero_red(Args a, int slot) {
  const int tube = blockIdx.x * blockDim.x + threadIdx.x;
  if (tube >= a.ntubes) return;
  unsigned long long* r = red_of(a, tube, slot);
#pragma unroll
  for (int k = 0; k < kRed; k++) r[k] = 0ull;
}

__device__ __forceinline__ double red_val(const unsigned long long* p) {
  return __longlong_as_double((long long)*p);
}

// dt and finish check of every running tube (src/vac.t:63-81):
// getdt_courant (src/vac.t:641-699), getdt_visc (src/vacusr.viscosity.t:509-541,
// dtdiffpar 0.25, maxviscoef reset), then timetofinish (src/vac.t:585-605).
// A tube whose dt is not a positive finite number is stopped as FAILED.
__global__ void kb_dt(Args a) {
  const int tube = blockIdx.x * blockDim.x + threadIdx.x;
  if (tube >= a.ntubes || a.state[tube] != RUNNING) return;
  const unsigned long long* r = red_of(a, tube, SAC_SLOT_W);
  double cm = 0.0;
  for (int d = 0; d < a.ndim; d++) cm = fmax(cm, red_val(r + a.ndim + d));
  double dt = 1.0e99;
  if (cm > 1.0e-99) dt = fmin(dt, a.courantpar / cm);
  if (a.visc_on) {
    const double mv = red_val(a.visc + tube);
    for (int d = 0; d < a.ndim; d++) {
      const double dxm = a.dxmin[2 * tube + d];
      dt = fmin(dt, 0.25 / (mv / (dxm * dxm)));
    }
    a.visc[tube] = 0ull;
  }
  if (a.it[tube] >= a.itmax || a.t[tube] >= a.tmax) {
    a.state[tube] = FINISHED;
  } else if (dt < a.dtmin) {
    a.state[tube] = STOPPED;
  } else if (!(dt > 0.0) || !isfinite(dt)) {
    a.state[tube] = FAILED;
  } else {
    a.dt[tube] = dt;
  }
}

// t += dt, it += 1 after the step; counts the tubes still running
__global__ void kb_tick(Args a) {
  const int tube = blockIdx.x * blockDim.x + threadIdx.x;
  if (tube >= a.ntubes || a.state[tube] != RUNNING) return;
  a.t[tube] += a.dt[tube];
  a.it[tube] += 1;
  atomicAdd(a.running, 1u);
}

__global__ void kb_zero_u32(unsigned int* p) { *p = 0u; }

}  // namespace batch
}  // namespace sac

namespace {

using namespace sac::batch;

thread_local std::string g_berr;

int bfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_berr = buf;
  return code;
}

#define BCHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return bfail(SAC_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// launch shapes the autotuner chooses from.  Tubes: tiles of tb consecutive
// cells of one tube, one cell per thread (kb_tube_stage<tb>).  Patches: one cell per thread, tiles of tbx x (tb/tbx)
// cells.  Every shape gives the same result bit for bit.
struct Shape {
  int ck, tb, tbx;
};
const Shape kShapes1[] = {{1, 64, 0}, {1, 128, 0}, {1, 256, 0}};
const Shape kShapes2[] = {{1, 64, 16},  {1, 64, 32},  {1, 64, 64},   {1, 128, 16},
                          {1, 128, 32}, {1, 128, 64}, {1, 128, 128}, {1, 256, 16},
                          {1, 256, 32}, {1, 256, 64}, {1, 256, 128}, {1, 256, 256}};

struct Batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int ndim = 1;
  int ntubes = 0;
  std::vector<int> nxy;          // [tube][2] mesh cells along x and y (1 along y for tubes)
  std::vector<int> pitch;        // row pitch of each tube's grid
  std::vector<long long> start;  // packed grid origins
  std::vector<long long> dxo;    // offsets of each tube's cell widths
  std::vector<int> active;       // tubes the tiles cover (running ones after a compaction)
  long long fs = 0;
  int max_frame = 0;             // ghost cells of the largest patch
  Args a{};
  // device buffers
  long long* d_start = nullptr;
  int* d_nxy = nullptr;
  int* d_pitch = nullptr;
  long long* d_dxo = nullptr;
  Tile* d_tiles = nullptr;
  Tile* d_ltiles[2] = {nullptr, nullptr};
  int* d_active = nullptr;
  double* d_reg[kSlots] = {nullptr, nullptr, nullptr, nullptr};
  double* d_bg = nullptr;
  double* d_dx = nullptr;
  double* d_R = nullptr;
  double* d_T = nullptr;
  double* d_save = nullptr;  // state kept aside while the autotuner runs
  unsigned long long* d_red = nullptr;
  unsigned long long* d_visc = nullptr;
  double* d_t = nullptr;
  double* d_dt = nullptr;
  int* d_it = nullptr;
  int* d_state = nullptr;
  double* d_dxmin = nullptr;
  unsigned int* d_running = nullptr;
  Shape shape{2, 128, 0};
  int ntiles = 0;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
  int compactions = 0;  // tile rebuilds over the running tubes so far
  long long bytes = 0;
  long long steps_run = 0;

  int gx(int i) const { return nxy[2 * i] + 4; }
  int gy(int i) const { return ndim == 2 ? nxy[2 * i + 1] + 4 : 1; }
};

int balloc(Batch& b, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return bfail(SAC_E_NOMEM, "hipMalloc of %zu bytes: %s", bytes,
                                    hipGetErrorString(e));
  b.bytes += (long long)bytes;
  e = hipMemset(*p, 0, bytes);
  if (e != hipSuccess) return bfail(SAC_E_HIP, "hipMemset: %s", hipGetErrorString(e));
  return 0;
}

// patch counts at or below which the line tiles use short segments
constexpr size_t kFewPatches = 64;

int set_shape(Batch& b, Shape s) {
  // the graph's kernel nodes hold the tile tables freed below: destroy the
  // executable graph first (the stream is idle here), so no graph ever
  // refers to freed memory
  if (b.graph) {
    hipGraphExecDestroy(b.graph);
    b.graph = nullptr;
  }
  b.graph_steps = 0;
  std::vector<Tile> tiles;
  if (b.ndim == 1) {
    const int span = s.ck * s.tb;
    for (int i : b.active)
      for (int c = 0; c < b.nxy[2 * i]; c += span) tiles.push_back({i, 2 + c, 0});
  } else {
    const int tby = s.tb / s.tbx;
    for (int i : b.active)
      for (int c1 = 0; c1 < b.nxy[2 * i + 1]; c1 += tby)
        for (int c0 = 0; c0 < b.nxy[2 * i]; c0 += s.tbx) tiles.push_back({i, 2 + c0, 2 + c1});
  }
  if (tiles.empty()) tiles.push_back({b.active.empty() ? 0 : b.active[0], 2, 2});
  if (b.ndim == 2) {
    // line tiles of each axis: every logical block of each running patch's
    // grid.  Whole lines while many patches run; once few remain (their
    // steps are latency-bound), segments of 32 cells give more, shorter
    // blocks.  The result does not depend on the segment length.
    const bool few = b.active.size() <= kFewPatches;
    for (int K = 0; K < 2; K++) {
      std::vector<Tile> lt;
      for (int i : b.active) {
        sac::GridDesc g{};
        g.ndim = 2;
        g.ndir = NDIR;
        g.G[0] = b.gx(i);
        g.G[1] = b.gy(i);
        g.G[2] = 1;
        g.Mlo[0] = g.Mlo[1] = 2;
        g.Mhi[0] = g.G[0] - 3;
        g.Mhi[1] = g.G[1] - 3;
        const int ext = g.Mhi[K] - g.Mlo[K] + 1;
        const int seg = few ? std::min(ext, 32) : ext;
        const long long nb = (K == 0) ? sac::line::nblocks<2, 0, double>(g, seg)
                                      : sac::line::nblocks<2, 1, double>(g, seg);
        for (long long k = 0; k < nb; k++) lt.push_back({i, (int)k, seg});
      }
      if (lt.empty()) lt.push_back({b.active.empty() ? 0 : b.active[0], 1 << 30, 1});
      if (b.d_ltiles[K]) hipFree(b.d_ltiles[K]);
      b.d_ltiles[K] = nullptr;
      BCHK(hipMalloc((void**)&b.d_ltiles[K], sizeof(Tile) * lt.size()));
      BCHK(hipMemcpy(b.d_ltiles[K], lt.data(), sizeof(Tile) * lt.size(), hipMemcpyHostToDevice));
      b.a.ltiles[K] = b.d_ltiles[K];
      b.a.nltiles[K] = (int)lt.size();
    }
  }
  if (b.d_tiles) hipFree(b.d_tiles);
  b.d_tiles = nullptr;
  BCHK(hipMalloc((void**)&b.d_tiles, sizeof(Tile) * tiles.size()));
  BCHK(hipMemcpy(b.d_tiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice));
  b.ntiles = (int)tiles.size();
  b.shape = s;
  b.a.tiles = b.d_tiles;
  b.a.ntiles = b.ntiles;
  b.a.tbx = s.tbx;
  BCHK(hipMemcpy(b.d_active, b.active.data(), sizeof(int) * b.active.size(),
                 hipMemcpyHostToDevice));
  b.a.active = b.d_active;
  b.a.nactive = (int)b.active.size();
  return 0;
}

void enqueue_ghost(Batch& b, double* f, int nfield) {
  if (b.active.empty()) return;
  if (b.ndim == 1) {
    const dim3 gg((unsigned)((4LL * b.active.size() + 255) / 256));
    hipLaunchKernelGGL(kb_ghost, gg, dim3(256), 0, b.stream, b.a, f, nfield);
  } else {
    const dim3 gg((unsigned)((b.max_frame + 255) / 256), (unsigned)b.active.size());
    hipLaunchKernelGGL(kb_ghost2, gg, dim3(256), 0, b.stream, b.a, f, nfield);
  }
}

// one tube stage kernel at the batch's tile length
template <bool RK4L>
void launch_tube_stage(Batch& b, dim3 gt, dim3 bt, int wct, int base, int out, double da) {
  const Args& a = b.a;
  if (b.shape.tb == 64)
    hipLaunchKernelGGL((kb_tube_stage<64, RK4L>), gt, bt, 0, b.stream, a, wct, base, out, da);
  else if (b.shape.tb == 128)
    hipLaunchKernelGGL((kb_tube_stage<128, RK4L>), gt, bt, 0, b.stream, a, wct, base, out, da);
  else
    hipLaunchKernelGGL((kb_tube_stage<256, RK4L>), gt, bt, 0, b.stream, a, wct, base, out, da);
}

template <int NDIM, int CK>
void enqueue_step_ck(Batch& b) {
  const Args& a = b.a;
  hipStream_t st = b.stream;
  const dim3 gt(b.ntiles), bt(b.shape.tb);
  const dim3 gn((b.ntubes + 255) / 256), bn(256);
  const int W = SAC_SLOT_W, S1 = SAC_SLOT_S1, S2 = SAC_SLOT_S2, S3 = SAC_SLOT_S3;
  hipLaunchKernelGGL(kb_dt, gn, bn, 0, st, a);
  // (wct, base, out, da, rk4_last) of the four RK4 stages, as in the grid
  // path: S1=W+dt/2 k1; S2=W+dt/2 k2; S3=W+dt k3;
  // W=(W+dt/6 k4)+(S1+2 S2+S3-4 W)/3
  struct St {
    int wct, base, out;
    double da;
    int last;
  };
  const St stages[4] = {{W, W, S1, 2.0, 0},
                        {S1, W, S2, 2.0, 0},
                        {S2, W, S3, 1.0, 0},
                        {S3, W, W, 6.0, 1}};
  for (const St& s : stages) {
    if constexpr (NDIM == 1) {
      hipLaunchKernelGGL(kb_zero_red, gn, bn, 0, st, a, s.out);
      if (s.last) launch_tube_stage<true>(b, gt, bt, s.wct, s.base, s.out, s.da);
      else launch_tube_stage<false>(b, gt, bt, s.wct, s.base, s.out, s.da);
      continue;
    } else {
      hipLaunchKernelGGL(kb_pline<0>, dim3(a.nltiles[0]), dim3(sac::line::Tile<0, double>::T), 0, st,
                         a, s.wct, s.out);
      hipLaunchKernelGGL(kb_pline<1>, dim3(a.nltiles[1]), dim3(sac::line::Tile<1, double>::T), 0, st,
                         a, s.wct, s.out);
    }
    if (s.last)
      hipLaunchKernelGGL((kb_final<NDIM, CK, true>), gt, bt, 0, st, a, s.wct, s.base, s.out, s.da);
    else
      hipLaunchKernelGGL((kb_final<NDIM, CK, false>), gt, bt, 0, st, a, s.wct, s.base, s.out,
                         s.da);
  }
  hipLaunchKernelGGL(kb_tick, gn, bn, 0, st, a);
}

void enqueue_step(Batch& b) {
  if (b.ndim == 2)
    enqueue_step_ck<2, 1>(b);
  else
    enqueue_step_ck<1, 1>(b);
}

void enqueue_cmax(Batch& b, int slot) {
  const dim3 gn((b.ntubes + 255) / 256), bn(256);
  hipLaunchKernelGGL(kb_zero_red, gn, bn, 0, b.stream, b.a, slot);
  const dim3 gt(b.ntiles), bt(b.shape.tb);
  if (b.ndim == 2)
    hipLaunchKernelGGL((kb_cmax<2, 1>), gt, bt, 0, b.stream, b.a, slot);
  else
    hipLaunchKernelGGL((kb_cmax<1, 1>), gt, bt, 0, b.stream, b.a, slot);
}

// capture `steps` time steps (and the running-tube counter reset) in one graph
int build_graph(Batch& b, int steps) {
  if (b.graph) {
    hipGraphExecDestroy(b.graph);
    b.graph = nullptr;
  }
  hipGraph_t g = nullptr;
  BCHK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
  // the running-tube counter is cleared before the last step only, so after
  // a replay it holds the number of tubes that took that step (kb_tick)
  for (int s = 0; s < steps; s++) {
    if (s == steps - 1) hipLaunchKernelGGL(kb_zero_u32, dim3(1), dim3(1), 0, b.stream, b.a.running);
    enqueue_step(b);
  }
  BCHK(hipStreamEndCapture(b.stream, &g));
  hipError_t e = hipGraphInstantiate(&b.graph, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return bfail(SAC_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  b.graph_steps = steps;
  return 0;
}

// per-tube run state (t, dt, it, state, maxviscoef), kept aside by the autotuner
struct RunState {
  std::vector<double> t, dt;
  std::vector<int> it, state;
  std::vector<unsigned long long> visc;
};

int save_state(Batch& b, RunState& r) {
  const size_t n = b.ntubes;
  r.t.resize(n);
  r.dt.resize(n);
  r.it.resize(n);
  r.state.resize(n);
  r.visc.resize(n);
  BCHK(hipMemcpy(r.t.data(), b.d_t, sizeof(double) * n, hipMemcpyDeviceToHost));
  BCHK(hipMemcpy(r.dt.data(), b.d_dt, sizeof(double) * n, hipMemcpyDeviceToHost));
  BCHK(hipMemcpy(r.it.data(), b.d_it, sizeof(int) * n, hipMemcpyDeviceToHost));
  BCHK(hipMemcpy(r.state.data(), b.d_state, sizeof(int) * n, hipMemcpyDeviceToHost));
  BCHK(hipMemcpy(r.visc.data(), b.d_visc, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost));
  return 0;
}

int restore_state(Batch& b, const RunState& r) {
  const size_t n = b.ntubes;
  BCHK(hipMemcpy(b.d_t, r.t.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  BCHK(hipMemcpy(b.d_dt, r.dt.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  BCHK(hipMemcpy(b.d_it, r.it.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  BCHK(hipMemcpy(b.d_state, r.state.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  BCHK(hipMemcpy(b.d_visc, r.visc.data(), sizeof(unsigned long long) * n, hipMemcpyHostToDevice));
  return 0;
}

// drop finished tubes from the tiles (their blocks would only exit), then
// recapture the graph; the arithmetic of the running tubes is unchanged
int compact(Batch& b, int steps_per_graph) {
  std::vector<int> st(b.ntubes);
  BCHK(hipMemcpy(st.data(), b.d_state, sizeof(int) * b.ntubes, hipMemcpyDeviceToHost));
  std::vector<int> act;
  for (int i : b.active)
    if (st[i] == RUNNING) act.push_back(i);
  if (act.empty() || act.size() == b.active.size()) return 0;
  b.active.swap(act);
  if (int rc = set_shape(b, b.shape)) return rc;
  b.compactions += 1;
  // graph runs recapture; direct-launch runs (steps_per_graph <= 0) never capture
  return steps_per_graph > 0 ? build_graph(b, steps_per_graph) : 0;
}

}  // namespace

extern "C" {

const char* sac_batch_last_error(void) { return g_berr.c_str(); }

int sac_batch_init_nd(int ndim, int ntubes, const int* ncell, int nw, const double* w,
                      const double* dx, const double* eqpar, int neqpar, int nu_,
                      double courantpar, double tmax, int itmax, sac_batch_handle* out) {
  g_berr.clear();
  if (!out) return bfail(SAC_E_ARG, "out is null");
  *out = nullptr;
  if (ndim != 1 && ndim != 2)
    return bfail(SAC_E_UNSUPPORTED, "ensembles are of 1-D tubes or 2-D patches (ndim=%d)", ndim);
  if (ntubes < 1 || !ncell || !w || !dx || !eqpar) return bfail(SAC_E_ARG, "bad arguments");
  if (ndim == 2 && ntubes > 65535)
    return bfail(SAC_E_ARG, "at most 65535 patches per ensemble (%d)", ntubes);
  if (nw != NE + NBG) return bfail(SAC_E_UNSUPPORTED, "tubes carry nw=%d variables (ndir=3)", NE + NBG);
  if (neqpar < 2 || eqpar[0] <= 0.0) return bfail(SAC_E_ARG, "gamma must be > 0");
  if (eqpar[1] != 0.0) return bfail(SAC_E_UNSUPPORTED, "resistive MHD (eta /= 0) is not built");
  if (!(courantpar > 0.0) || !(tmax > 0.0) || itmax < 1)
    return bfail(SAC_E_ARG, "courantpar, tmax and itmax must be positive");
  for (int i = 0; i < ntubes; i++)
    for (int d = 0; d < ndim; d++)
      if (ncell[ndim * i + d] < 6)
        return bfail(SAC_E_ARG, "tube %d has %d < 6 cells along axis %d", i, ncell[ndim * i + d],
                     d + 1);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return bfail(SAC_E_HIP, "no HIP device available");
  Batch* b = new Batch();
  auto bail = [&](int code) {
    sac_batch_free(b);
    return code;
  };
  hipGetDevice(&b->device);
  b->ndim = ndim;
  b->ntubes = ntubes;
  b->nxy.assign(2 * (size_t)ntubes, 1);
  for (int i = 0; i < ntubes; i++)
    for (int d = 0; d < ndim; d++) b->nxy[2 * i + d] = ncell[ndim * i + d];
  b->start.resize(ntubes);
  b->pitch.resize(ntubes);
  b->dxo.resize(ntubes);
  long long off = 0, doff = 0;
  for (int i = 0; i < ntubes; i++) {
    const int G0 = b->gx(i), G1 = b->gy(i);
    b->start[i] = off;  // each tube starts on a 128-byte line
    b->pitch[i] = (ndim == 1) ? G0 : (G0 + 15) / 16 * 16;
    off += ((long long)b->pitch[i] * G1 + 15) / 16 * 16;
    b->dxo[i] = doff;
    doff += G0 + (ndim == 2 ? G1 : 0);
    b->max_frame = std::max(b->max_frame, 4 * G0 + 4 * (G1 - 4));
  }
  b->fs = (off + 63) / 64 * 64;
  if (hipStreamCreate(&b->stream) != hipSuccess) return bail(bfail(SAC_E_HIP, "hipStreamCreate"));
  int rc = 0;
  const size_t fb = sizeof(double) * (size_t)b->fs;
  for (int s = 0; s < kSlots; s++)
    if ((rc = balloc(*b, (void**)&b->d_reg[s], fb * NE))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_bg, fb * NBG))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_R, fb * NE))) return bail(rc);
  if (ndim == 2 && (rc = balloc(*b, (void**)&b->d_T, fb * NT2))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_dx, sizeof(double) * (size_t)doff))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_start, sizeof(long long) * ntubes))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_nxy, sizeof(int) * 2 * ntubes))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_pitch, sizeof(int) * ntubes))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_dxo, sizeof(long long) * ntubes))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_red, sizeof(unsigned long long) * ntubes * kSlots * kRed)))
    return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_visc, sizeof(unsigned long long) * ntubes))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_t, sizeof(double) * ntubes))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_dt, sizeof(double) * ntubes))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_it, sizeof(int) * ntubes))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_state, sizeof(int) * ntubes))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_dxmin, sizeof(double) * 2 * ntubes))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_running, sizeof(unsigned int)))) return bail(rc);
  if ((rc = balloc(*b, (void**)&b->d_active, sizeof(int) * ntubes))) return bail(rc);
  b->active.resize(ntubes);
  for (int i = 0; i < ntubes; i++) b->active[i] = i;
  // host layout per tube: w(G0[,G1], nw) column-major, tubes one after
  // another; dx: G0 values (then G1 values) per tube
  std::vector<double> pk((size_t)b->fs * (NE + NBG), 0.0);
  std::vector<double> dxmin(2 * (size_t)ntubes, 1.0);
  long long hoff = 0;
  for (int i = 0; i < ntubes; i++) {
    const int G0 = b->gx(i), G1 = b->gy(i);
    for (int v = 0; v < NE + NBG; v++)
      for (int r = 0; r < G1; r++)
        memcpy(&pk[(size_t)v * b->fs + b->start[i] + (size_t)r * b->pitch[i]],
               w + hoff + ((size_t)v * G1 + r) * G0, sizeof(double) * G0);
    for (int d = 0; d < ndim; d++) {
      const double* dd = dx + b->dxo[i] + (d == 1 ? G0 : 0);
      const int n = b->nxy[2 * i + d];
      double mn = 1.0e99;
      for (int c = 2; c <= n + 1; c++) {
        if (!(dd[c] > 0.0)) return bail(bfail(SAC_E_ARG, "dx must be > 0 (tube %d)", i));
        mn = std::min(mn, dd[c]);
      }
      dxmin[2 * i + d] = mn;
    }
    hoff += (long long)G0 * G1 * (NE + NBG);
  }
  if (hipMemcpy(b->d_reg[SAC_SLOT_W], pk.data(), fb * NE, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_bg, pk.data() + (size_t)b->fs * NE, fb * NBG, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_dx, dx, sizeof(double) * (size_t)doff, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_start, b->start.data(), sizeof(long long) * ntubes, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_nxy, b->nxy.data(), sizeof(int) * 2 * ntubes, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_pitch, b->pitch.data(), sizeof(int) * ntubes, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_dxo, b->dxo.data(), sizeof(long long) * ntubes, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_dxmin, dxmin.data(), sizeof(double) * 2 * ntubes, hipMemcpyHostToDevice) !=
          hipSuccess)
    return bail(bfail(SAC_E_HIP, "upload failed"));
  Args& a = b->a;
  a.ndim = ndim;
  a.ntubes = ntubes;
  a.fs = b->fs;
  a.start = b->d_start;
  a.nxy = b->d_nxy;
  a.pitch = b->d_pitch;
  a.dxo = b->d_dxo;
  for (int s = 0; s < kSlots; s++) a.reg[s] = b->d_reg[s];
  a.bg = b->d_bg;
  a.dx = b->d_dx;
  a.R = b->d_R;
  a.T = b->d_T;
  a.red = b->d_red;
  a.visc = b->d_visc;
  a.t = b->d_t;
  a.dt = b->d_dt;
  a.it = b->d_it;
  a.state = b->d_state;
  a.dxmin = b->d_dxmin;
  a.gamma = eqpar[0];
  const double nu = (nu_ >= 1 && nu_ <= neqpar) ? eqpar[nu_ - 1] : 0.0;
  a.visc_on = std::fabs(nu) > 1.0e-99;
  a.courantpar = courantpar;
  a.tmax = tmax;
  a.dtmin = 1.0e-99;
  a.itmax = itmax;
  a.running = b->d_running;
  if ((rc = set_shape(*b, ndim == 1 ? Shape{1, 256, 0} : Shape{1, 256, 64}))) return bail(rc);
  // initial getboundary (background and state) and cmax of W
  enqueue_ghost(*b, b->d_bg, NBG);
  enqueue_ghost(*b, b->d_reg[SAC_SLOT_W], NE);
  enqueue_cmax(*b, SAC_SLOT_W);
  if (hipStreamSynchronize(b->stream) != hipSuccess)
    return bail(bfail(SAC_E_HIP, "initial fill: %s", hipGetErrorString(hipGetLastError())));
  *out = b;
  return 0;
}

int sac_batch_init(int ntubes, const int* ncell, int nw, const double* w, const double* dx,
                   const double* eqpar, int neqpar, int nu_, double courantpar, double tmax,
                   int itmax, sac_batch_handle* out) {
  return sac_batch_init_nd(1, ntubes, ncell, nw, w, dx, eqpar, neqpar, nu_, courantpar, tmax,
                           itmax, out);
}

int sac_batch_free(sac_batch_handle h) {
  if (!h) return 0;
  Batch* b = (Batch*)h;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->graph) hipGraphExecDestroy(b->graph);
  for (int s = 0; s < kSlots; s++) hipFree(b->d_reg[s]);
  hipFree(b->d_bg);
  hipFree(b->d_dx);
  hipFree(b->d_R);
  hipFree(b->d_T);
  hipFree(b->d_save);
  hipFree(b->d_start);
  hipFree(b->d_nxy);
  hipFree(b->d_pitch);
  hipFree(b->d_dxo);
  hipFree(b->d_tiles);
  hipFree(b->d_ltiles[0]);
  hipFree(b->d_ltiles[1]);
  hipFree(b->d_active);
  hipFree(b->d_red);
  hipFree(b->d_visc);
  hipFree(b->d_t);
  hipFree(b->d_dt);
  hipFree(b->d_it);
  hipFree(b->d_state);
  hipFree(b->d_dxmin);
  hipFree(b->d_running);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int sac_batch_tune(sac_batch_handle h, int trial_steps, int* ck, int* tb, double* ms_per_step) {
  if (!h) return bfail(SAC_E_ARG, "null handle");
  Batch& b = *(Batch*)h;
  if (trial_steps < 1) trial_steps = 2;
  // keep the state, the reductions and the run counters aside
  const size_t fb = sizeof(double) * (size_t)b.fs * NE;
  if (!b.d_save) {
    int rc = balloc(b, (void**)&b.d_save, fb);
    if (rc) return rc;
  }
  std::vector<unsigned long long> red0((size_t)b.ntubes * kSlots * kRed);
  RunState r0;
  BCHK(hipMemcpy(b.d_save, b.d_reg[SAC_SLOT_W], fb, hipMemcpyDeviceToDevice));
  BCHK(hipMemcpy(red0.data(), b.d_red, sizeof(unsigned long long) * red0.size(),
                 hipMemcpyDeviceToHost));
  if (int rc = save_state(b, r0)) return rc;
  double best = 1e300;
  Shape bs = b.shape;
  const Shape* shapes = (b.ndim == 1) ? kShapes1 : kShapes2;
  const int nshape = (b.ndim == 1) ? (int)(sizeof(kShapes1) / sizeof(Shape))
                                   : (int)(sizeof(kShapes2) / sizeof(Shape));
  for (int k = 0; k < nshape; k++) {
    const Shape& s = shapes[k];
    if (int rc = set_shape(b, s)) return rc;
    if (int rc = build_graph(b, trial_steps)) return rc;
    BCHK(hipGraphLaunch(b.graph, b.stream));  // warm-up replay
    BCHK(hipStreamSynchronize(b.stream));
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, b.stream);
    BCHK(hipGraphLaunch(b.graph, b.stream));
    hipEventRecord(e1, b.stream);
    BCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    const double per = ms / trial_steps;
    if (per < best) {
      best = per;
      bs = s;
    }
    // restore the state for the next candidate
    BCHK(hipMemcpy(b.d_reg[SAC_SLOT_W], b.d_save, fb, hipMemcpyDeviceToDevice));
    BCHK(hipMemcpy(b.d_red, red0.data(), sizeof(unsigned long long) * red0.size(),
                   hipMemcpyHostToDevice));
    if (int rc = restore_state(b, r0)) return rc;
  }
  if (int rc = set_shape(b, bs)) return rc;
  if (ck) *ck = bs.ck;
  if (tb) *tb = bs.tb;
  if (ms_per_step) *ms_per_step = best;
  return 0;
}

int sac_batch_set_shape(sac_batch_handle h, int index) {
  if (!h) return bfail(SAC_E_ARG, "null handle");
  Batch& b = *(Batch*)h;
  const Shape* shapes = (b.ndim == 1) ? kShapes1 : kShapes2;
  const int nshape = (b.ndim == 1) ? (int)(sizeof(kShapes1) / sizeof(Shape))
                                   : (int)(sizeof(kShapes2) / sizeof(Shape));
  if (index < 0 || index >= nshape)
    return bfail(SAC_E_ARG, "shape index %d outside 0..%d", index, nshape - 1);
  return set_shape(b, shapes[index]);
}

int sac_batch_shape(sac_batch_handle h, int* ck, int* tbx, int* tby) {
  if (!h) return bfail(SAC_E_ARG, "null handle");
  const Batch& b = *(const Batch*)h;
  if (ck) *ck = b.shape.ck;
  if (tbx) *tbx = (b.ndim == 1) ? b.shape.tb * b.shape.ck : b.shape.tbx;
  if (tby) *tby = (b.ndim == 1) ? 1 : b.shape.tb / b.shape.tbx;
  return 0;
}

int sac_batch_run(sac_batch_handle h, int steps_per_graph, long long max_steps,
                  long long* steps_done) {
  if (!h) return bfail(SAC_E_ARG, "null handle");
  Batch& b = *(Batch*)h;
  if (steps_per_graph < 0) {
    // direct launches, one step at a time (for tools that trace launches)
    long long done = 0;
    while (done < max_steps) {
      hipLaunchKernelGGL(kb_zero_u32, dim3(1), dim3(1), 0, b.stream, b.a.running);
      enqueue_step(b);
      unsigned int running = 0;
      BCHK(hipMemcpyAsync(&running, b.d_running, sizeof(running), hipMemcpyDeviceToHost,
                          b.stream));
      BCHK(hipStreamSynchronize(b.stream));
      done += 1;
      if (running == 0) break;
      if (2 * (size_t)running <= b.active.size()) {
        if (int rc = compact(b, -1)) return rc;
      }
    }
    b.steps_run += done;
    if (steps_done) *steps_done = done;
    return 0;
  }
  if (steps_per_graph < 1) steps_per_graph = 32;
  if (!b.graph || b.graph_steps != steps_per_graph) {
    if (int rc = build_graph(b, steps_per_graph)) return rc;
  }
  long long done = 0;
  while (done < max_steps) {
    BCHK(hipGraphLaunch(b.graph, b.stream));
    unsigned int running = 0;
    BCHK(hipMemcpyAsync(&running, b.d_running, sizeof(running), hipMemcpyDeviceToHost, b.stream));
    BCHK(hipStreamSynchronize(b.stream));
    done += steps_per_graph;
    if (running == 0) break;
    // once at most half of the covered tubes run, rebuild the tiles
    if (2 * (size_t)running <= b.active.size()) {
      if (int rc = compact(b, steps_per_graph)) return rc;
    }
  }
  b.steps_run += done;
  if (steps_done) *steps_done = done;
  return 0;
}

int sac_batch_status(sac_batch_handle h, double* t, int* it, int* state) {
  if (!h) return bfail(SAC_E_ARG, "null handle");
  Batch& b = *(Batch*)h;
  if (t) BCHK(hipMemcpy(t, b.d_t, sizeof(double) * b.ntubes, hipMemcpyDeviceToHost));
  if (it) BCHK(hipMemcpy(it, b.d_it, sizeof(int) * b.ntubes, hipMemcpyDeviceToHost));
  if (state) BCHK(hipMemcpy(state, b.d_state, sizeof(int) * b.ntubes, hipMemcpyDeviceToHost));
  return 0;
}

int sac_batch_download(sac_batch_handle h, double* w) {
  if (!h || !w) return bfail(SAC_E_ARG, "null argument");
  Batch& b = *(Batch*)h;
  BCHK(hipStreamSynchronize(b.stream));
  const size_t fb = sizeof(double) * (size_t)b.fs;
  std::vector<double> pk((size_t)b.fs * (NE + NBG));
  BCHK(hipMemcpy(pk.data(), b.d_reg[SAC_SLOT_W], fb * NE, hipMemcpyDeviceToHost));
  BCHK(hipMemcpy(pk.data() + (size_t)b.fs * NE, b.d_bg, fb * NBG, hipMemcpyDeviceToHost));
  long long hoff = 0;
  for (int i = 0; i < b.ntubes; i++) {
    const int G0 = b.gx(i), G1 = b.gy(i);
    for (int v = 0; v < NE + NBG; v++)
      for (int r = 0; r < G1; r++)
        memcpy(w + hoff + ((size_t)v * G1 + r) * G0,
               &pk[(size_t)v * b.fs + b.start[i] + (size_t)r * b.pitch[i]], sizeof(double) * G0);
    hoff += (long long)G0 * G1 * (NE + NBG);
  }
  return 0;
}

int sac_batch_memory(sac_batch_handle h, long long* bytes) {
  if (!h) return bfail(SAC_E_ARG, "null handle");
  if (bytes) *bytes = ((Batch*)h)->bytes;
  return 0;
}

int sac_batch_compactions(sac_batch_handle h, int* compactions, int* covered) {
  if (!h) return bfail(SAC_E_ARG, "null handle");
  const Batch& b = *(const Batch*)h;
  if (compactions) *compactions = b.compactions;
  if (covered) *covered = (int)b.active.size();
  return 0;
}

}  // extern "C"
