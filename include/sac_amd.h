/*
 * sac_amd.h -- C ABI of the MI355X (gfx950) implementation of SAC's
 * explicit perturbation-MHD update.
 *
 * The library owns device copies of the state; the caller owns the host
 * array.  Host arrays use the reference layout: the Fortran array
 * w(ixG1,ixG2,ixG3,nw), column-major, index 1 fastest, two ghost layers per
 * active axis (src/vacdef.t:18-23), variables in the order of
 * src/vacpar.t.mhd:9-13 (rho, m1..m_ndir, e, b1..b_ndir, eb, rhob,
 * bg1..bg_ndir).
 *
 * Every function returns 0 on success or a positive code (SAC_E_*); the
 * message of the last failure on the calling thread is returned by
 * sac_dev_last_error().  There is no CPU fallback: without a usable GPU,
 * sac_dev_init fails with SAC_E_HIP.
 *
 * Plain C: pointers and sizes only, all arguments by value or pointer, so a
 * Fortran host binds them with ISO_C_BINDING interfaces (host/sac_bind.f90).
 */
#ifndef SAC_AMD_H
#define SAC_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define SAC_OK 0
#define SAC_E_ARG 1         /* invalid argument */
#define SAC_E_UNSUPPORTED 2 /* configuration outside this library's scope */
#define SAC_E_HIP 3         /* HIP runtime error (no device, launch failure) */
#define SAC_E_NOMEM 4       /* device allocation failed */

/* boundary codes for typeB (src/vacgrid.t:257-464) */
#define SAC_BC_CONT 1
#define SAC_BC_FIXED 2
#define SAC_BC_PERIODIC 3
#define SAC_BC_SYMM 4
#define SAC_BC_SYMM0 5
#define SAC_BC_ASYMM 6
#define SAC_BC_CONT1 7   /* first-order extrapolation (src/vacgrid.t:376-407) */
#define SAC_BC_FIXED1 8  /* extrapolated once, then restored */
#define SAC_BC_GRAD1 9   /* extrapolated once, then stored difference + edge value */
#define SAC_BC_CONTCD4 10 /* mirror about the edge cell (src/vacgrid.t:260-290) */
#define SAC_BC_ZERO 11    /* contCD4's copy in primitive variables, the whole grid
                             converted before and after (src/vacgrid.t:294-345) */

/* typeadvance codes (src/vac.t:406-480) */
#define SAC_ADV_ONESTEP 1
#define SAC_ADV_TWOSTEP 2   /* with a 'cd4' predictor; typepred1='nul' is onestep */
#define SAC_ADV_THREESTEP 3
#define SAC_ADV_FOURSTEP 4
#define SAC_ADV_STERCK 5
#define SAC_ADV_JAMESON 6

/* precision of the evolving fields */
#define SAC_PREC_FP64 0  /* everything in FP64 */
#define SAC_PREC_MIXED 1 /* evolving fields, fluxes and RK registers in FP32 with compensated
                            flux differences; background fields, pressure recovery from the
                            split energy, register update, cmax/dt and maxviscoef in FP64 */

/* device registers addressed by the slot-level entry points */
#define SAC_SLOT_W 0  /* the state w */
#define SAC_SLOT_S1 1 /* stage registers */
#define SAC_SLOT_S2 2
#define SAC_SLOT_S3 3 /* RK4's third stage register */

/* kernels reported by sac_dev_timing_get */
#define SAC_K_FINAL 0 /* cross terms, gravity, register update, cmax reductions */
#define SAC_K_AXIS 1  /* per-axis fluxes, setnu and hyperdiffusion (ndim launches) */
#define SAC_K_GHOST 2
#define SAC_K_CMAX 3
#define SAC_K_COPY 4
#define SAC_K_AXIS1 5 /* k_axis launches of axis 1, 2, 3 separately */
#define SAC_K_AXIS2 6
#define SAC_K_AXIS3 7
#define SAC_K_CLOSURE 8 /* learned-closure inference launches (one per axis) */
#define SAC_K_PLANE 9   /* 3-D grids: the x and y work of z-plane tiles (one launch per stage) */

typedef void *sac_handle;

/*
 * Set up one problem on the current device.  Replaces the grid, boundary and
 * method set-up of readparameters/boundsetup/gridsetup1
 * (src/vacio.t:5-530, src/vacgrid.t:9-160,556-661) for the 'cd4' path.
 *   ndim, ndir      dimensionality (1<=ndim<=ndir<=3)
 *   nx[ndim]        mesh cells per axis (>= 6)
 *   nw              number of variables, must be 4+3*ndir
 *   iws[niws]       1-based variables with typefull='cd4' (iw_full,
 *                   src/vacio.t:249-262); must be the evolving set 1..2+2*ndir
 *   eqpar[neqpar]   equation + user parameters (gamma_=1, eta_=2,
 *                   gravity d at grav0_+d, hyperdiffusion switch at nu_);
 *                   eta must be 0
 *   dx              cell widths per axis after gridsetup1, concatenated:
 *                   dx_1[ixG1], dx_2[ixG2], dx_3[ixG3] (tensor grid)
 *   typeB[iw*2*ndim+iB]  boundary code per variable and boundary region
 *                   (regions ordered lower/upper per axis, src/vacgrid.t:53-69)
 *   fullgridini     1 if the ini file carried ghost cells
 *   typeadvance     SAC_ADV_* (onestep, twostep, threestep, fourstep, sterck,
 *                   jameson of src/vac.t:406-475)
 *   precision       SAC_PREC_FP64 or SAC_PREC_MIXED (the host arrays stay FP64
 *                   either way; FP32 storage is converted at upload/download)
 */
int sac_dev_init(int ndim, int ndir, const int *nx, int nw, const int *iws, int niws,
                 const double *eqpar, int neqpar, int grav0_, int nu_, const double *dx,
                 const int *typeB, int fullgridini, int typeadvance, int precision,
                 sac_handle *out);

/* Release all device memory of a handle. */
int sac_dev_free(sac_handle h);

/* Copy the state w (reference layout, nw variables, whole grid) to / from
 * the device.  The static background (variables eb..bg) is taken from the
 * upload and returned unchanged by the download; the ghost cells of an
 * uploaded background are filled again by the next sac_dev_getboundary. */
int sac_dev_upload(sac_handle h, const double *w);
int sac_dev_download(sac_handle h, double *w);

/* Asynchronous saves (the D2H of the state at save times, src/vac.t:56-58 ->
 * savefile, src/vacio.t:901-958).  sac_dev_save_begin takes a device copy of
 * the state on the handle's stream and starts its copy to pinned host memory
 * on a second stream; later calls (sac_dev_advance ...) may run at once and
 * overlap that copy.  sac_dev_save_ready reports (without waiting) whether
 * the copy has finished; sac_dev_save_end waits for it and writes w exactly
 * as sac_dev_download would have at the time of sac_dev_save_begin.  One
 * save at a time per handle.  The device copy and the pinned buffer (one
 * state each) are allocated at the first save and kept until sac_dev_free. */
int sac_dev_save_begin(sac_handle h);
int sac_dev_save_ready(sac_handle h, int *ready);
int sac_dev_save_end(sac_handle h, double *w);

/* Slot-level copies of the evolving variables (1..2+2*ndir) for stage-level
 * parity checks; the background part of w is ignored on upload and filled
 * from the device background on download. */
int sac_dev_upload_slot(sac_handle h, int slot, const double *w);
int sac_dev_download_slot(sac_handle h, int slot, double *w);

/* getboundary(t,1,nw,1,ndim,w) on the state (src/vacgrid.t:195-512); the
 * first call fills the background's ghosts and stores the 'fixed' boundary
 * values as the reference does (later calls fill the background's ghosts
 * only after an upload). */
int sac_dev_getboundary(sac_handle h);

/* getdt_courant (src/vac.t:641-699): dt = courantpar/max(cmax_d/dx_d) over
 * the mesh; also returns courantmax.  The max comes from the reduction fused
 * into the last stage that produced the state. */
int sac_dev_getdt_courant(sac_handle h, double courantpar, double *dt, double *courantmax);

/* sac_dev_getdt_courant without a handle and without any device allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid it).  The caller passes device pointers in the reference layout
 * (unpadded, index 1 fastest, two ghost layers per active axis):
 *   w    the 2+2*ndir evolving fields (double, or float for SAC_PREC_MIXED)
 *   bg   the 2+ndir background fields eb, rhob, bg1.. (double)
 *   dx   the cell widths dx_1[ixG1], dx_2[ixG2], dx_3[ixG3] (double)
 * and the stream to use (a hipStream_t, or null for the null stream).  The
 * grid may hold at most 2^36 cells.  The reduction slots are a static device
 * array of 6 uint64 per GPU, kept until the process ends: calls that
 * overlap on one GPU race.  dt and courantmax are returned on the host, as
 * by sac_dev_getdt_courant, and equal its values on the same state. */
int sac_getdt_courant_noalloc(int ndim, int ndir, const int *nx, int precision, const void *w,
                              const double *bg, const double *dx, double gamma,
                              double courantpar, void *stream, double *dt, double *courantmax);

/* getdt_special of the gravity+hyperdiffusion user module
 * (src/vacusr.f:1403-1421): dt = min(dt, getdt_visc, getdt_grav) with
 * getdt_visc (src/vacusr.viscosity.t:509-541) using the maxviscoef gathered
 * since the previous call, which it then resets, and getdt_grav
 * (src/vacusr.gravity.t:97-138). */
int sac_dev_getdt_special(sac_handle h, double *dt);

/* advance(iw_full,w) (src/vac.t:224-281) with dimsplit=F, sourcesplit=F:
 * one full Runge-Kutta step of the state including every getboundary. */
int sac_dev_advance(sac_handle h, double dt);

/* Kernel schedule of the stages of 3-D grids: with on=1 (the default) the x
 * and y work of a stage runs in one pass over 2-D tiles of each z-plane
 * (the x-y cross terms included), then the z line tiles; with on=0 one
 * line-tile pass per axis.  Both evaluate the same terms of centdiff4,
 * setnu and addsource_visc (src/vaccd.t:7-78, src/vacusr.viscosity.t:2-455)
 * and agree to rounding (the cross terms are summed in another order).  The
 * learned closure and Powell's sources use the per-axis passes.  No
 * counterpart in the reference (a schedule choice of this library). */
int sac_dev_set_plane_tiles(sac_handle h, int on);

/* divbfix: Powell's div B source terms (addsource_divb,
 * src/vacphys.mhd0.t:256-297) added to every advect1 when on and ndim > 1
 * (the reference adds nothing in 1-D).  Off after sac_dev_init. */
int sac_dev_set_divbfix(sac_handle h, int on);

/* Resistive MHD with a constant eta = eqpar(eta_) > 0 (set at sac_dev_init;
 * eta < 0 needs the user's specialeta and is rejected there): compactres
 * selects addsource_res1 (compact) over addsource_res2
 * (src/vacphys.mhdres.t:96-301); dtdiffpar is the par value the diffusion
 * limit dtdiffpar*min(dx)^2/eta uses until getdt_visc assigns 0.25 to it
 * (getdt_res, src/vacphys.mhdres.t:5-44).  Defaults: res2, 0.5. */
int sac_dev_set_resistive(sac_handle h, int compactres, double dtdiffpar);

/* advect1 with method 'cd4' (src/vac.t:490-549) on device registers:
 * slot w <- w + qdt*R(slot wCT), then getboundary(w). */
int sac_dev_advect1(sac_handle h, double qdt, int wct_slot, int w_slot);

/* Register copy of the evolving variables, whole grid: the reference's
 * w1(ixG^S,1:nw)=w(ixG^S,1:nw) (src/vac.t:403,444,447).  Done by a plain
 * streaming kernel, which also serves as the known-size copy that HBM
 * counters are calibrated against. */
int sac_dev_copy_slot(sac_handle h, int dst_slot, int src_slot);

/* maxviscoef (src/vacdef.t:176) accumulated by setnu since the last
 * getdt_special. */
int sac_dev_get_maxviscoef(sac_handle h, double *v);
int sac_dev_set_maxviscoef(sac_handle h, double v);

/* Kernel timing with HIP events on the library's stream. */
int sac_dev_timing(sac_handle h, int enable);
int sac_dev_timing_get(sac_handle h, int kernel, double *total_ms, long long *count);
int sac_dev_synchronize(sac_handle h);

/* Bytes of device memory held by the handle. */
int sac_dev_memory(sac_handle h, long long *bytes);

/* Message of the last failure on this thread ("" if none). */
const char *sac_dev_last_error(void);

/* ------------------------------------------------------------------------
 * Learned hyperdiffusion closure (FP64 handles).  A per-face MLP
 * (21 -> 64 -> 64 -> 64 -> 7, BF16 matrix-core arithmetic with FP32
 * accumulation) predicts setnu's ratio md3/md1 (src/vacusr.viscosity.t:
 * 413-427) of every variable of an axis from a 7-point stencil of rho',
 * v_K and a transverse b, normalised by the local density and sound speed.
 * Faces where the shock sensor (largest normalised jump of rho' or v_K over
 * the stencil) reaches the threshold, or near the grid edge, keep the exact
 * ratio.  No reference counterpart: it replaces the coefficient computation
 * inside setnu when enabled; everything else of the step is unchanged.
 * ---------------------------------------------------------------------- */
int sac_closure_nparam(void);
/* Parameters (He-normal from seed), Adam state, room for `capacity`
 * training samples and per-block gradients for up to max_blocks wavefronts. */
int sac_closure_init(sac_handle h, long long capacity, int max_blocks, unsigned seed);
int sac_closure_set_params(sac_handle h, const float *p, int n);
int sac_closure_get_params(sac_handle h, float *p, int n);
/* Append (stencil, exact ratio) samples of every mesh face of every axis of
 * register `slot`; faces past the sensor threshold get weight 0. */
int sac_closure_collect(sac_handle h, int slot, long long *total);
/* The same samples from register `slot` of another handle `src` (any grid
 * and dimensionality, FP64 storage, same device) into h's buffer: a
 * network trained on one configuration's samples from the runs of others. */
int sac_closure_collect_from(sac_handle h, sac_handle src, int slot, long long *total);
int sac_closure_clear(sac_handle h);
/* `steps` Adam steps of blocks*64 samples each; sample k of the run is
 * ((base + k) * stride) mod n (stride coprime with n walks a permutation).
 * mean_loss: mean squared error of the last step. */
int sac_closure_train(sac_handle h, int steps, int blocks, double lr, long long base,
                      long long stride, double *mean_loss);
/* Loss and parameter gradient of one forward/backward pass over the given
 * samples (features [n][32] FP32, rounded to BF16; targets [n][8]: 7 ratios/6
 * (<0: none) and a weight); n a multiple of 64. */
int sac_closure_grad(sac_handle h, const float *feat, const float *targ, int nsamples,
                     float *grad, double *loss);
/* Use the learned ratios in advance/advect1 (on=1), stop using them (0), or
 * (2) route the exact ratios through the same storage, a check of the
 * learned path's indexing that must follow the exact run to FP32 rounding. */
int sac_closure_enable(sac_handle h, int on, double sensor);

/* ------------------------------------------------------------------------
 * Ensembles of independent 1-D tubes or 2-D patches (ndim=1 or 2, ndir=3,
 * nw=13, 'cont' boundaries, 'fourstep', no gravity).  No reference
 * counterpart as one call: each tube is one serial run of the reference's
 * main loop (src/vac.t:55-106) with its own dt (getdt_courant +
 * getdt_visc), t and it.  The ensemble advances on the device in hipGraph
 * replays of several steps.  A tube stops at t >= tmax or it >= itmax
 * (state 1), at dt < dtmin = 1e-99 (state 3, the reference's third finish
 * condition, reached when a tube blows up), or when dt is not a number
 * (state 2).
 * ---------------------------------------------------------------------- */
typedef void *sac_batch_handle;

/* ncell[ntubes] mesh cells per tube (>= 6).  w: per tube the reference array
 * w(ncell+4, nw) column-major, tubes one after another.  dx: per tube ncell+4
 * cell widths after gridsetup1.  eqpar as for sac_dev_init (gamma at 1, eta
 * at 2 must be 0, hyperdiffusion switch at nu_). */
int sac_batch_init(int ntubes, const int *ncell, int nw, const double *w, const double *dx,
                   const double *eqpar, int neqpar, int nu_, double courantpar, double tmax,
                   int itmax, sac_batch_handle *out);
/* The same for tubes (ndim=1) or patches (ndim=2): ncell[ntubes*ndim] mesh
 * cells per tube and axis (>= 6 each); w per tube w(G1[,G2], nw)
 * column-major (G = ncell+4); dx per tube the G1 widths along x, then (2-D)
 * the G2 widths along y.  At most 65535 patches.  sac_batch_init is this
 * call with ndim=1. */
int sac_batch_init_nd(int ndim, int ntubes, const int *ncell, int nw, const double *w,
                      const double *dx, const double *eqpar, int neqpar, int nu_,
                      double courantpar, double tmax, int itmax, sac_batch_handle *out);
int sac_batch_free(sac_batch_handle h);
/* Autotuner: times trial_steps steps for every launch shape from the
 * current state, restores the state, keeps the fastest shape.  Tubes: LDS
 * line tiles of tb in {64,128,256} cells, one per thread (ck = 1).
 * Patches: one cell per thread, tiles of tbx x tb/tbx cells for tb in
 * {64,128,256} and tbx in {16,...,tb} (12 shapes). */
int sac_batch_tune(sac_batch_handle h, int trial_steps, int *ck, int *tb, double *ms_per_step);
/* Use launch shape `index` of the autotuner's list (tubes 0..2, patches
 * 0..11) without timing. */
int sac_batch_set_shape(sac_batch_handle h, int index);
/* The current tile: cells per thread (ck) and the tile's extent along x and
 * y in cells (tubes: tbx = threads*ck, tby = 1; patches: ck = 1). */
int sac_batch_shape(sac_batch_handle h, int *ck, int *tbx, int *tby);
/* Advance every tube to its end, replaying a graph of steps_per_graph steps
 * at most max_steps steps in total; returns the steps issued.
 * steps_per_graph < 0 launches the kernels of each step directly instead
 * (same results; for tools that trace kernel launches).  Once at most
 * half of the tubes the tiles cover still run, the tiles and the graph are
 * rebuilt over the running ones (the results do not change). */
int sac_batch_run(sac_batch_handle h, int steps_per_graph, long long max_steps,
                  long long *steps_done);
/* Per tube: t, it and state (0 running, 1 finished, 2 failed, 3 dt < dtmin). */
int sac_batch_status(sac_batch_handle h, double *t, int *it, int *state);
/* The ensemble state in the layout of sac_batch_init(_nd)'s w. */
int sac_batch_download(sac_batch_handle h, double *w);
int sac_batch_memory(sac_batch_handle h, long long *bytes);
/* Tile rebuilds so far (a run rebuilds the tiles over the running tubes
 * once at most half of the covered ones still run) and the number of tubes
 * the tiles cover now.  No counterpart in the reference (its runs are one
 * tube per process); for callers that check the compaction schedule. */
int sac_batch_compactions(sac_batch_handle h, int *compactions, int *covered);
const char *sac_batch_last_error(void);


#ifdef __cplusplus
}
#endif
#endif
