/*
 * sac_oracle.c -- CPU restatement of SAC's explicit perturbation-MHD update.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * Every routine restates the reference Fortran (file:line given with each
 * function) statement by statement: same index ranges, same left-to-right
 * evaluation order, and persistent zero-initialised scratch arrays in place
 * of the reference's COMMON tmp/tmp2 and its static (-fno-automatic) local
 * arrays.  Index conventions follow the reference: 1-based, ixGlo=1,
 * ghost width dixB=2, mesh ixM = 3..nx+2, grid ixG = 1..nx+4; unused
 * dimensions have extent 1.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   gradient4         closed form: exact for polynomials of degree <= 4,
 *                     zero on the two outermost mesh layers.
 *   gradient1/1L/1R   closed form on linear data; edge zeroing pattern.
 *   getflux+centdiff4 against a textbook ideal-MHD flux function of the
 *                     total state (uniform magnetic background, arbitrary
 *                     density/energy background): split form == textbook.
 *   getptotal_bg and the background energy terms: same textbook check.
 *   getcmax           B=0 gives c_s+|v|; fast-magnetosonic closed form.
 *   setnu             hand-computed known answers for a step profile,
 *                     zero for constant and linear profiles, boundedness
 *                     nu <= 6 c_tot c_hyp dx, mirror-layer cells.
 *   addsource_visc    conservation of sum(rho) for interior data, mirror
 *                     and axis-permutation symmetry, hand-computed step.
 *   addsource_grav    closed-form solution of a uniform state under gravity.
 *   getboundary       cont/periodic/fixed semantics on index-coded data.
 *   RK4 (fourstep)    equals the textbook Butcher form applied to the
 *                     oracle's own right-hand side.
 *   whole step        hydrostatic zero-perturbation stays exactly zero,
 *                     uniform state stays uniform, axis permutation.
 *   getdt_courant     courantpar / max(cmax_d/dx_d) with a textbook fast
 *                     speed, on random and uniform states.
 *   getdt_visc        0.25 min(dx)^2/maxviscoef for set values, the reset to
 *                     0, no limit when nothing was collected; maxviscoef is
 *                     the largest setnu coefficient (known-answer step); the
 *                     first step of a run has only the courant limit and the
 *                     second uses the value the first collected.
 *   getdt_grav        1/sqrt(|g|/min(dx)) at it == itmin, kept afterwards,
 *                     absent when the first call comes after itmin.
 * The order of getboundary calls within a step is restated from the source
 * text and pinned through the invariants and the integrator checks above.
 */
#include "sac_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NWMAX 13
#define SMALLDOUBLE 1.0e-99 /* src/vacdef.t:50 */
#define BIGDOUBLE 1.0e99

typedef struct {
  int lo[3], hi[3];
} box;

struct sac_oracle {
  int ndim, ndir, nw;
  int G[3];          /* ixGhi (ixGlo = 1) */
  int Mlo[3], Mhi[3];
  long n;            /* cells of the grid */
  int E[3];          /* extents of the setnu tmp_nuI array (ixG+2) */
  int D[3];          /* extents of the setnu d-arrays (ixG+1) */
  long nE, nD;
  /* variable indices, 1-based (src/vacpar.t.mhd:9-13) */
  int rho_, m0_, e_, b0_, eb_, rhob_, bg0_;
  double eqpar[32];
  int neqpar, grav0_, nu_;
  int typeB[NWMAX][6];
  int ipairB[6];
  int fixedB[NWMAX][6];
  double *fixB; /* [6][nw][n] */
  int initialized;
  int fullgridini;
  int iws[NWMAX + 1];
  int niws;
  double maxviscoef;
  int it, itmin;
  double t;
  double dtgrav;
  double *x, *dx; /* [ndim][n] */
  double *w;      /* [nw][n] */
  /* persistent scratch */
  double *tmp, *tmp2;
  double *cd_v, *cd_f, *cd_fb;
  double *gf_fb;
  double *bd_p;     /* pressure of primitive() ('zero' boundaries) */
  double *csound2, *cfast2;
  double *gc_cmax;
  double *av_nushk; /* [3][n] */
  double *av_rhoL, *av_rhoR, *av_rhoC, *av_VL, *av_VR, *av_BL, *av_BR;
  double *av_L, *av_R, *av_C, *av_nuL, *av_nuR;
  double *sn_cmax, *sn_tmp_nu, *sn_tmp_nuI;
  double *sn_d1R, *sn_d1L, *sn_d3R, *sn_d3L;
  double *sn_md3R, *sn_md3L, *sn_md1R, *sn_md1L;
  double *dv_tmpdt, *dv_nushk;
  double *db_divb;
  double *rs_current; /* [3][n], components 1..3 (getcurrent's common array) */
  double *w1, *w2, *w3;
  int divbfix;
  int compactres;     /* resistive sources: addsource_res1 (T) or res2 (F) */
  double dtdiffpar;   /* common dtdiffpar: the par value until getdt_visc sets 0.25 */
};

/* ------------------------------------------------------------------ */
/* index helpers                                                        */

static inline long IDX(const sac_oracle *o, int i1, int i2, int i3) {
  return (long)(i1 - 1) + (long)o->G[0] * ((long)(i2 - 1) + (long)o->G[1] * (long)(i3 - 1));
}
static inline long IDXE(const sac_oracle *o, int i1, int i2, int i3) {
  return (long)(i1 - 1) + (long)o->E[0] * ((long)(i2 - 1) + (long)o->E[1] * (long)(i3 - 1));
}
static inline long IDXD(const sac_oracle *o, int i1, int i2, int i3) {
  return (long)(i1 - 1) + (long)o->D[0] * ((long)(i2 - 1) + (long)o->D[1] * (long)(i3 - 1));
}

#define FOR_BOX(b)                                         \
  for (int i3 = (b).lo[2]; i3 <= (b).hi[2]; i3++)          \
    for (int i2 = (b).lo[1]; i2 <= (b).hi[1]; i2++)        \
      for (int i1 = (b).lo[0]; i1 <= (b).hi[0]; i1++)

#define C(a) ((a)[IDX(o, i1, i2, i3)])
#define F(a, iw) ((a)[(long)((iw)-1) * o->n + IDX(o, i1, i2, i3)])
/* access shifted by s along idim */
#define SH(a, idim, s) \
  ((a)[IDX(o, i1 + ((idim) == 1) * (s), i2 + ((idim) == 2) * (s), i3 + ((idim) == 3) * (s))])
#define FSH(a, iw, idim, s)                      \
  ((a)[(long)((iw)-1) * o->n +                   \
       IDX(o, i1 + ((idim) == 1) * (s), i2 + ((idim) == 2) * (s), i3 + ((idim) == 3) * (s))])

static box grid_box(const sac_oracle *o) {
  box b;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = 1;
    b.hi[d] = o->G[d];
  }
  return b;
}
static box mesh_box(const sac_oracle *o) {
  box b;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = o->Mlo[d];
    b.hi[d] = o->Mhi[d];
  }
  return b;
}
/* ^LADDn on every active dimension */
static box box_add(const sac_oracle *o, box b, int n) {
  for (int d = 0; d < o->ndim; d++) {
    b.lo[d] -= n;
    b.hi[d] += n;
  }
  return b;
}
/* ^LADD n*kr(idim,^D): widen along one axis (1-based idim) */
static box box_add_dir(box b, int idim, int n) {
  b.lo[idim - 1] -= n;
  b.hi[idim - 1] += n;
  return b;
}

static double *zalloc(long n) {
  double *p = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  return p;
}

static double eq_gamma(const sac_oracle *o) { return o->eqpar[0]; }
static double eq_grav(const sac_oracle *o, int idim) {
  int slot = o->grav0_ + idim; /* 1-based */
  if (slot < 1 || slot > o->neqpar) return 0.0;
  return o->eqpar[slot - 1];
}
static double eq_nu(const sac_oracle *o) {
  if (o->nu_ < 1 || o->nu_ > o->neqpar) return 0.0;
  return o->eqpar[o->nu_ - 1];
}

/* ------------------------------------------------------------------ */
/* construction                                                         */

sac_oracle *oracle_new(int ndim, int ndir, const int *nx, const double *eqpar,
                       int neqpar, int grav0_, int nu_, const int *typeB,
                       int fullgridini) {
  if (ndim < 1 || ndim > 3 || ndir < ndim || ndir > 3 || neqpar < 2 || neqpar > 32) return NULL;
  sac_oracle *o = (sac_oracle *)calloc(1, sizeof(sac_oracle));
  o->ndim = ndim;
  o->ndir = ndir;
  o->rho_ = 1;
  o->m0_ = 1;
  o->e_ = ndir + 2;
  o->b0_ = o->e_;
  o->eb_ = o->b0_ + ndir + 1;
  o->rhob_ = o->eb_ + 1;
  o->bg0_ = o->rhob_;
  o->nw = o->bg0_ + ndir;
  o->n = 1;
  o->nE = 1;
  o->nD = 1;
  for (int d = 0; d < 3; d++) {
    if (d < ndim) {
      if (nx[d] < 6) {
        free(o);
        return NULL;
      }
      o->G[d] = nx[d] + 4;
      o->Mlo[d] = 3;
      o->Mhi[d] = nx[d] + 2;
      o->E[d] = o->G[d] + 2;
      o->D[d] = o->G[d] + 1;
    } else {
      o->G[d] = 1;
      o->Mlo[d] = o->Mhi[d] = 1;
      o->E[d] = 1;
      o->D[d] = 1;
    }
    o->n *= o->G[d];
    o->nE *= o->E[d];
    o->nD *= o->D[d];
  }
  memcpy(o->eqpar, eqpar, sizeof(double) * (size_t)neqpar);
  o->neqpar = neqpar;
  o->grav0_ = grav0_;
  o->nu_ = nu_;
  for (int iw = 0; iw < o->nw; iw++)
    for (int iB = 0; iB < 2 * ndim; iB++) o->typeB[iw][iB] = typeB[iw * 2 * ndim + iB];
  /* periodic pairs (src/vacgrid.t:109-120): default regions pair lower/upper */
  for (int iB = 0; iB < 2 * ndim; iB++) {
    o->ipairB[iB] = -1;
    if (o->typeB[0][iB] == OR_BC_PERIODIC) {
      int jB = (iB % 2 == 0) ? iB + 1 : iB - 1;
      if (o->typeB[0][jB] != OR_BC_PERIODIC) {
        free(o);
        return NULL; /* 'Error in BoundSetup: No periodic pair' */
      }
      o->ipairB[iB] = jB;
    }
  }
  o->fullgridini = fullgridini;
  o->niws = 0;
  for (int iw = 1; iw <= o->e_ + ndir; iw++) o->iws[o->niws++] = iw;
  o->dtgrav = BIGDOUBLE;
  long n = o->n;
  o->x = zalloc(3 * n);
  o->dx = zalloc(3 * n);
  o->w = zalloc(o->nw * n);
  o->fixB = zalloc(6L * o->nw * n);
  o->tmp = zalloc(n);
  o->tmp2 = zalloc(n);
  o->cd_v = zalloc(n);
  o->cd_f = zalloc(n);
  o->cd_fb = zalloc(n);
  o->gf_fb = zalloc(n);
  o->bd_p = zalloc(n);
  o->csound2 = zalloc(n);
  o->cfast2 = zalloc(n);
  o->gc_cmax = zalloc(n);
  o->av_nushk = zalloc(3 * n);
  o->av_rhoL = zalloc(n);
  o->av_rhoR = zalloc(n);
  o->av_rhoC = zalloc(n);
  o->av_VL = zalloc(n);
  o->av_VR = zalloc(n);
  o->av_BL = zalloc(n);
  o->av_BR = zalloc(n);
  o->av_L = zalloc(n);
  o->av_R = zalloc(n);
  o->av_C = zalloc(n);
  o->av_nuL = zalloc(n);
  o->av_nuR = zalloc(n);
  o->sn_cmax = zalloc(n);
  o->sn_tmp_nu = zalloc(n);
  o->sn_tmp_nuI = zalloc(o->nE);
  o->sn_d1R = zalloc(o->nD);
  o->sn_d1L = zalloc(o->nD);
  o->sn_d3R = zalloc(o->nD);
  o->sn_d3L = zalloc(o->nD);
  o->sn_md3R = zalloc(n);
  o->sn_md3L = zalloc(n);
  o->sn_md1R = zalloc(n);
  o->sn_md1L = zalloc(n);
  o->dv_tmpdt = zalloc(n);
  o->dv_nushk = zalloc(3 * n);
  o->db_divb = zalloc(n);
  o->rs_current = zalloc(3 * n);
  o->dtdiffpar = 0.5; /* par/DEFAULT */
  o->w1 = zalloc(o->nw * n);
  o->w2 = zalloc(o->nw * n);
  o->w3 = zalloc(o->nw * n);
  return o;
}

void oracle_free(sac_oracle *o) {
  if (!o) return;
  double *ps[] = {o->x, o->dx, o->w, o->fixB, o->tmp, o->tmp2, o->cd_v, o->cd_f, o->cd_fb,
                  o->gf_fb, o->bd_p, o->csound2, o->cfast2, o->gc_cmax, o->av_nushk, o->av_rhoL,
                  o->av_rhoR, o->av_rhoC, o->av_VL, o->av_VR, o->av_BL, o->av_BR, o->av_L,
                  o->av_R, o->av_C, o->av_nuL, o->av_nuR, o->sn_cmax, o->sn_tmp_nu,
                  o->sn_tmp_nuI, o->sn_d1R, o->sn_d1L, o->sn_d3R, o->sn_d3L, o->sn_md3R,
                  o->sn_md3L, o->sn_md1R, o->sn_md1L, o->dv_tmpdt, o->dv_nushk, o->db_divb,
                  o->rs_current, o->w1, o->w2, o->w3};
  for (size_t i = 0; i < sizeof(ps) / sizeof(ps[0]); i++) free(ps[i]);
  free(o);
}

int oracle_nw(const sac_oracle *o) { return o->nw; }
long oracle_ncell(const sac_oracle *o) { return o->n; }

void oracle_poison_scratch(sac_oracle *o, double value) {
  long n = o->n;
  double *ps[] = {o->tmp, o->tmp2, o->cd_v, o->cd_f, o->cd_fb, o->gf_fb, o->csound2,
                  o->cfast2, o->gc_cmax, o->av_rhoL, o->av_rhoR, o->av_rhoC, o->av_VL,
                  o->av_VR, o->av_BL, o->av_BR, o->av_L, o->av_R, o->av_C, o->sn_cmax,
                  o->sn_tmp_nu, o->sn_md3R, o->sn_md3L, o->sn_md1R, o->sn_md1L, o->dv_tmpdt};
  for (size_t k = 0; k < sizeof(ps) / sizeof(ps[0]); k++)
    for (long i = 0; i < n; i++) ps[k][i] = value;
  for (long i = 0; i < o->nE; i++) o->sn_tmp_nuI[i] = value;
  for (long i = 0; i < o->nD; i++) {
    o->sn_d1R[i] = value;
    o->sn_d1L[i] = value;
    o->sn_d3R[i] = value;
    o->sn_d3L[i] = value;
  }
  /* av_nuL/av_nuR/av_nushk are deliberately left alone: outside the mesh the
   * reference reads their never-written (zero) entries, see addsource_visc */
}

int oracle_set_iws(sac_oracle *o, const int *iws, int niws) {
  if (niws < 0 || niws > NWMAX) return -1;
  for (int i = 0; i < niws; i++) {
    if (iws[i] < 1 || iws[i] > o->nw) return -1;
    o->iws[i] = iws[i];
  }
  o->niws = niws;
  return 0;
}

int oracle_set_w(sac_oracle *o, const double *w) {
  memcpy(o->w, w, sizeof(double) * (size_t)(o->nw * o->n));
  return 0;
}
int oracle_get_w(const sac_oracle *o, double *w) {
  memcpy(w, o->w, sizeof(double) * (size_t)(o->nw * o->n));
  return 0;
}

/* gridsetup1 (src/vacgrid.t:556-629): linear extrapolation of x into the
 * ghost layers (and one extra layer), then dx = (x_{i+1}-x_{i-1})/2 */
int oracle_set_x(sac_oracle *o, const double *x) {
  int nd = o->ndim;
  int Q[3]; /* qx(ixG^LL^LADD1:): indices 0..G+1 on active axes */
  long nq = 1;
  for (int d = 0; d < 3; d++) {
    Q[d] = (d < nd) ? o->G[d] + 2 : 1;
    nq *= Q[d];
  }
  double *qx = zalloc(nq * nd);
  /* j = (j1,j2,j3) in reference indices on active axes; inactive axes 1 */
#define QI(j1, j2, j3)                                                 \
  ((long)((nd > 0) ? (j1) : 0) +                                       \
   (long)Q[0] * ((long)((nd > 1) ? (j2) : 0) + (long)Q[1] * (long)((nd > 2) ? (j3) : 0)))
  box g = grid_box(o);
  for (int jd = 0; jd < nd; jd++) FOR_BOX(g) {
      qx[jd * nq + QI(i1, i2, i3)] = x[(long)jd * o->n + IDX(o, i1, i2, i3)];
    }
  for (int idim = 1; idim <= nd; idim++) {
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = (d < nd) ? 0 : 1;
      hi[d] = (d < nd) ? o->G[d] + 1 : 1;
    }
    for (int side = 1; side >= 0; side--) { /* upper layers first */
      int mn, mx, ixe, ixf;
      if (side == 1) {
        mx = o->G[idim - 1] + 1;
        mn = o->fullgridini ? o->G[idim - 1] + 1 : o->Mhi[idim - 1] + 1;
        ixe = mn - 1;
        ixf = ixe - 1;
      } else {
        mn = 0;
        mx = o->fullgridini ? 0 : o->Mlo[idim - 1] - 1;
        ixe = mx + 1;
        ixf = ixe + 1;
      }
      for (int jd = 0; jd < nd; jd++)
        for (int ix = mn; ix <= mx; ix++) {
          int l2[3] = {lo[0], lo[1], lo[2]}, h2[3] = {hi[0], hi[1], hi[2]};
          l2[idim - 1] = h2[idim - 1] = ix;
          for (int j3 = l2[2]; j3 <= h2[2]; j3++)
            for (int j2 = l2[1]; j2 <= h2[1]; j2++)
              for (int j1 = l2[0]; j1 <= h2[0]; j1++) {
                int e[3] = {j1, j2, j3}, f[3] = {j1, j2, j3};
                e[idim - 1] = ixe;
                f[idim - 1] = ixf;
                int a = abs(ixe - ix);
                qx[jd * nq + QI(j1, j2, j3)] =
                    (1 + a) * qx[jd * nq + QI(e[0], e[1], e[2])] - a * qx[jd * nq + QI(f[0], f[1], f[2])];
              }
        }
    }
  }
  /* x(ixG)=qx(ixG); dx(ixG,idim)=half*(qx(jx)-qx(hx)) */
  for (int jd = 0; jd < nd; jd++) FOR_BOX(g) {
      o->x[(long)jd * o->n + IDX(o, i1, i2, i3)] = qx[jd * nq + QI(i1, i2, i3)];
      long jp = QI(i1 + (jd == 0), i2 + (jd == 1), i3 + (jd == 2));
      long jm = QI(i1 - (jd == 0), i2 - (jd == 1), i3 - (jd == 2));
      o->dx[(long)jd * o->n + IDX(o, i1, i2, i3)] = 0.5 * (qx[jd * nq + jp] - qx[jd * nq + jm]);
    }
#undef QI
  free(qx);
  return 0;
}

int oracle_get_dx(const sac_oracle *o, double *dx) {
  memcpy(dx, o->dx, sizeof(double) * (size_t)(o->ndim * o->n));
  return 0;
}

#define DX(idim) (o->dx[(long)((idim)-1) * o->n + IDX(o, i1, i2, i3)])

/* ------------------------------------------------------------------ */
/* getboundary (src/vacgrid.t:195-512) for cont/fixed/periodic/symm/asymm,
 * cont1/fixed1/grad1, contCD4 and zero */

static void primitive(sac_oracle *o, double *w);
static void conserve(sac_oracle *o, double *w);

static void getboundary(sac_oracle *o, double *w) {
  int nd = o->ndim;
  for (int iB = 0; iB < 2 * nd; iB++) {
    int idim = iB / 2 + 1;
    int upper = iB % 2;
    box r = grid_box(o);
    if (upper) {
      r.lo[idim - 1] = o->Mhi[idim - 1] + 1;
      r.hi[idim - 1] = o->G[idim - 1];
    } else {
      r.lo[idim - 1] = 1;
      r.hi[idim - 1] = o->Mlo[idim - 1] - 1;
    }
    for (int iw = 1; iw <= o->nw; iw++) {
      int tb = o->typeB[iw - 1][iB];
      if (tb == OR_BC_PERIODIC) {
        /* w(ix^S)=w(ixpair^S), ixpair = pair region shifted by dixB */
        int shift = upper ? -(o->Mhi[idim - 1] - o->Mlo[idim - 1] + 1)
                          : (o->Mhi[idim - 1] - o->Mlo[idim - 1] + 1);
        FOR_BOX(r) { F(w, iw) = FSH(w, iw, idim, shift); }
      } else if (tb == OR_BC_CONT || tb == OR_BC_FIXED) {
        int ixe = upper ? r.lo[idim - 1] - 1 : r.hi[idim - 1] + 1;
        if (o->fixedB[iw - 1][iB]) {
          const double *fb = o->fixB + ((long)iB * o->nw + (iw - 1)) * o->n;
          FOR_BOX(r) { F(w, iw) = C(fb); }
        } else if (tb == OR_BC_CONT || !o->fullgridini) {
          FOR_BOX(r) {
            int ix = (idim == 1) ? i1 : (idim == 2) ? i2 : i3;
            F(w, iw) = FSH(w, iw, idim, ixe - ix);
          }
        }
      } else if (tb == OR_BC_CONT1 || tb == OR_BC_FIXED1 || tb == OR_BC_GRAD1) {
        /* first-order extrapolation from the edge and the inner edge
         * (src/vacgrid.t:376-407); fixed1 restores stored values, grad1
         * adds stored differences to the current edge value */
        int ixe = upper ? r.lo[idim - 1] - 1 : r.hi[idim - 1] + 1;
        int ixf = upper ? ixe - 1 : ixe + 1;
        if (o->fixedB[iw - 1][iB]) {
          const double *fb = o->fixB + ((long)iB * o->nw + (iw - 1)) * o->n;
          if (tb == OR_BC_GRAD1) {
            FOR_BOX(r) {
              int ix = (idim == 1) ? i1 : (idim == 2) ? i2 : i3;
              F(w, iw) = C(fb) + FSH(w, iw, idim, ixe - ix);
            }
          } else {
            FOR_BOX(r) { F(w, iw) = C(fb); }
          }
        } else if (tb == OR_BC_CONT1 || !o->fullgridini) {
          FOR_BOX(r) {
            int ix = (idim == 1) ? i1 : (idim == 2) ? i2 : i3;
            int k = abs(ix - ixe);
            F(w, iw) = (k + 1) * FSH(w, iw, idim, ixe - ix) - k * FSH(w, iw, idim, ixf - ix);
          }
        }
      } else if (tb == OR_BC_CONTCD4) {
        /* mirror about the edge cell for two ghost layers (src/vacgrid.t:260-290):
         * upper ghosts ixmin <- ixmin-2, ixmax <- ixmin-3; lower ghosts
         * ixmin <- ixmax+3, ixmax <- ixmax+2 */
        int lo = r.lo[idim - 1], hi = r.hi[idim - 1];
        int src_lo = upper ? lo - 2 : hi + 3, src_hi = upper ? lo - 3 : hi + 2;
        FOR_BOX(r) {
          int ix = (idim == 1) ? i1 : (idim == 2) ? i2 : i3;
          if (ix == lo) F(w, iw) = FSH(w, iw, idim, src_lo - ix);
          else if (ix == hi) F(w, iw) = FSH(w, iw, idim, src_hi - ix);
        }
      } else if (tb == OR_BC_ZERO) {
        /* the contCD4 copy of the two ghost layers, made in primitive
         * variables: the whole grid is converted to primitive variables
         * before and back to conservative ones after (src/vacgrid.t:294-345) */
        int lo = r.lo[idim - 1], hi = r.hi[idim - 1];
        int src_lo = upper ? lo - 2 : hi + 3, src_hi = upper ? lo - 3 : hi + 2;
        primitive(o, w);
        FOR_BOX(r) {
          int ix = (idim == 1) ? i1 : (idim == 2) ? i2 : i3;
          if (ix == lo) F(w, iw) = FSH(w, iw, idim, src_lo - ix);
          else if (ix == hi) F(w, iw) = FSH(w, iw, idim, src_hi - ix);
        }
        conserve(o, w);
      } else if (tb == OR_BC_SYMM || tb == OR_BC_SYMM0 || tb == OR_BC_ASYMM) {
        int ixe = upper ? 2 * r.lo[idim - 1] - 1 : 2 * r.hi[idim - 1] + 1;
        FOR_BOX(r) {
          int ix = (idim == 1) ? i1 : (idim == 2) ? i2 : i3;
          F(w, iw) = FSH(w, iw, idim, (ixe - ix) - ix);
        }
        if (tb == OR_BC_ASYMM) FOR_BOX(r) { F(w, iw) = -F(w, iw); }
      }
    }
  }
  if (!o->initialized) {
    o->initialized = 1;
    for (int iB = 0; iB < 2 * nd; iB++) {
      int idim = iB / 2 + 1;
      int upper = iB % 2;
      box r = grid_box(o);
      if (upper) {
        r.lo[idim - 1] = o->Mhi[idim - 1] + 1;
        r.hi[idim - 1] = o->G[idim - 1];
      } else {
        r.lo[idim - 1] = 1;
        r.hi[idim - 1] = o->Mlo[idim - 1] - 1;
      }
      for (int iw = 1; iw <= o->nw; iw++) {
        int tb = o->typeB[iw - 1][iB];
        if ((tb == OR_BC_FIXED || tb == OR_BC_FIXED1 || tb == OR_BC_GRAD1) &&
            !o->fixedB[iw - 1][iB]) {
          o->fixedB[iw - 1][iB] = 1;
          double *fb = o->fixB + ((long)iB * o->nw + (iw - 1)) * o->n;
          if (tb == OR_BC_GRAD1) {
            /* differences to the edge value (src/vacgrid.t:489-494) */
            int ixe = upper ? r.lo[idim - 1] - 1 : r.hi[idim - 1] + 1;
            FOR_BOX(r) {
              int ix = (idim == 1) ? i1 : (idim == 2) ? i2 : i3;
              C(fb) = F(w, iw) - FSH(w, iw, idim, ixe - ix);
            }
          } else {
            FOR_BOX(r) { C(fb) = F(w, iw); }
          }
        }
      }
    }
  }
}

void oracle_getboundary(sac_oracle *o) { getboundary(o, o->w); }
void oracle_getboundary_on(sac_oracle *o, double *w) { getboundary(o, w); }

/* ------------------------------------------------------------------ */
/* pointwise physics (src/vacphys0.t.mhd)                               */

#define RHOT(w) (F(w, o->rho_) + F(w, o->rhob_))

/* getv (src/vacphys0.t.mhd:55-73) */
static void getv(sac_oracle *o, const double *w, box ix, int idim, double *v) {
  FOR_BOX(ix) { C(v) = F(w, o->m0_ + idim) / RHOT(w); }
}

/* getpthermal (src/vacphys0.t.mhd:163-183) */
static void getpthermal(sac_oracle *o, const double *w, box ix, double *p) {
  double g = eq_gamma(o);
  FOR_BOX(ix) {
    double s = 0.0;
    for (int c = 1; c <= o->ndir; c++) s = (c == 1) ? F(w, o->m0_ + c) * F(w, o->m0_ + c) : s + F(w, o->m0_ + c) * F(w, o->m0_ + c);
    C(p) = 0.5 * s / RHOT(w);
    double sb = 0.0, sbb = 0.0;
    for (int c = 1; c <= o->ndir; c++) {
      double b = F(w, o->b0_ + c);
      sb = (c == 1) ? b * b : sb + b * b;
      sbb = (c == 1) ? b * F(w, o->bg0_ + c) : sbb + b * F(w, o->bg0_ + c);
    }
    C(p) = C(p) + 0.5 * sb + sbb;
    C(p) = (g - 1.0) * (F(w, o->e_) - C(p));
  }
}

/* primitive (src/vacphys0.t.mhd:31-53) over the whole grid: e <- the
 * thermal pressure, m <- m/rho_t */
static void primitive(sac_oracle *o, double *w) {
  box g = grid_box(o);
  double *p = o->bd_p;
  getpthermal(o, w, g, p);
  FOR_BOX(g) { F(w, o->e_) = C(p); }
  for (int c = 1; c <= o->ndir; c++) FOR_BOX(g) { F(w, o->m0_ + c) = F(w, o->m0_ + c) / RHOT(w); }
}

/* conserve (src/vacphys0.t.mhd:5-28) over the whole grid:
 * e <- p/(gamma-1) + half*(rho_t*sum v^2 + sum b^2) + sum b*bg, m <- rho_t*v */
static void conserve(sac_oracle *o, double *w) {
  box g = grid_box(o);
  double gm1 = eq_gamma(o) - 1.0;
  FOR_BOX(g) {
    double sv = 0.0, sb = 0.0, sbb = 0.0;
    for (int c = 1; c <= o->ndir; c++) {
      double v = F(w, o->m0_ + c), b = F(w, o->b0_ + c);
      sv = (c == 1) ? v * v : sv + v * v;
      sb = (c == 1) ? b * b : sb + b * b;
      sbb = (c == 1) ? b * F(w, o->bg0_ + c) : sbb + b * F(w, o->bg0_ + c);
    }
    F(w, o->e_) = F(w, o->e_) / gm1 + 0.5 * (RHOT(w) * sv + sb) + sbb;
  }
  for (int c = 1; c <= o->ndir; c++) FOR_BOX(g) { F(w, o->m0_ + c) = RHOT(w) * F(w, o->m0_ + c); }
}

/* getcsound2 (src/vacphys0.t.mhd:137-160) */
static void getcsound2(sac_oracle *o, const double *w, box ix, double *cs2) {
  double g = eq_gamma(o);
  getpthermal(o, w, ix, cs2);
  FOR_BOX(ix) {
    double sbg = 0.0;
    for (int c = 1; c <= o->ndir; c++) {
      double bg = F(w, o->bg0_ + c);
      sbg = (c == 1) ? bg * bg : sbg + bg * bg;
    }
    C(cs2) = g * (C(cs2) + (g - 1.0) * (F(w, o->eb_) - 0.5 * sbg)) / RHOT(w);
  }
}

/* getcmax (src/vacphys0.t.mhd:77-112); csound2/cfast2 are SAVEd */
static void getcmax(sac_oracle *o, int *new_cmax, const double *w, box ix, int idim,
                    double *cmax) {
  if (*new_cmax) {
    *new_cmax = 0;
    getcsound2(o, w, ix, o->csound2);
    FOR_BOX(ix) {
      double s = 0.0;
      for (int c = 1; c <= o->ndir; c++) {
        double b = F(w, o->b0_ + c) + F(w, o->bg0_ + c);
        s = (c == 1) ? b * b : s + b * b;
      }
      C(o->cfast2) = s / RHOT(w) + C(o->csound2);
    }
  }
  FOR_BOX(ix) {
    double cf2 = C(o->cfast2), cs2 = C(o->csound2);
    double bd = F(w, o->b0_ + idim) + F(w, o->bg0_ + idim);
    C(cmax) = sqrt(0.5 * (cf2 + sqrt(cf2 * cf2 - 4.0 * cs2 * (bd * bd) / RHOT(w)))) +
              fabs(F(w, o->m0_ + idim) / RHOT(w));
  }
}

/* getptotal (src/vacphys0.t.mhd:186-205): perturbation total pressure */
static void getptotal(sac_oracle *o, const double *w, box ix, double *p) {
  double g = eq_gamma(o);
  FOR_BOX(ix) {
    double sbbg = 0.0, sb2 = 0.0, sm2 = 0.0;
    for (int c = 1; c <= o->ndir; c++) {
      double b = F(w, o->b0_ + c), bg = F(w, o->bg0_ + c), m = F(w, o->m0_ + c);
      sbbg = (c == 1) ? b * bg : sbbg + b * bg;
      sb2 = (c == 1) ? b * b : sb2 + b * b;
      sm2 = (c == 1) ? m * m : sm2 + m * m;
    }
    C(p) = (g - 2.0) * (sbbg + 0.5 * sb2);
    C(p) = (g - 1.0) * (F(w, o->e_) - 0.5 * sm2 / RHOT(w)) - C(p);
  }
}

/* getptotal_bg (src/vacphys0.t.mhd:208-224) */
static void getptotal_bg(sac_oracle *o, const double *w, box ix, double *p) {
  double g = eq_gamma(o);
  FOR_BOX(ix) {
    double s = 0.0;
    for (int c = 1; c <= o->ndir; c++) {
      double bg = F(w, o->bg0_ + c);
      s = (c == 1) ? bg * bg : s + bg * bg;
    }
    C(p) = (g - 1.0) * F(w, o->eb_) - 0.5 * (g - 2.0) * s;
  }
}

/* getflux (src/vacphys.mhd0.t:147-222) */
static void getflux(sac_oracle *o, const double *w, box ix, int iw, int idim, double *f,
                    int *transport) {
  *transport = 1;
  int nc = o->ndir;
  if (iw == o->rho_) {
    FOR_BOX(ix) { C(f) = F(w, o->rhob_) * F(w, o->m0_ + idim) / RHOT(w); }
    return;
  }
  if (iw > o->m0_ && iw <= o->m0_ + nc) {
    int c = iw - o->m0_;
    if (idim == c) {
      getptotal(o, w, ix, f);
      getptotal_bg(o, w, ix, o->gf_fb);
      FOR_BOX(ix) { C(o->gf_fb) = 0.0; }
      FOR_BOX(ix) {
        C(f) = C(f) + C(o->gf_fb) -
               (F(w, o->b0_ + c) * F(w, o->bg0_ + idim) + F(w, o->b0_ + idim) * F(w, o->bg0_ + c)) -
               F(w, o->b0_ + c) * F(w, o->b0_ + idim);
      }
    } else {
      FOR_BOX(ix) {
        C(f) = -(F(w, o->b0_ + c) * F(w, o->bg0_ + idim) + F(w, o->b0_ + idim) * F(w, o->bg0_ + c)) -
               F(w, o->b0_ + c) * F(w, o->b0_ + idim);
      }
    }
    return;
  }
  if (iw == o->e_) {
    getptotal(o, w, ix, f);
    getptotal_bg(o, w, ix, o->gf_fb);
    FOR_BOX(ix) { C(o->gf_fb) = 0.0; }
    FOR_BOX(ix) {
      double sbgm = 0.0, sbm = 0.0;
      for (int c = 1; c <= nc; c++) {
        double m = F(w, o->m0_ + c);
        sbgm = (c == 1) ? F(w, o->bg0_ + c) * m : sbgm + F(w, o->bg0_ + c) * m;
        sbm = (c == 1) ? F(w, o->b0_ + c) * m : sbm + F(w, o->b0_ + c) * m;
      }
      double rt = RHOT(w);
      C(f) = (F(w, o->m0_ + idim) * (C(f) + C(o->gf_fb)) - F(w, o->b0_ + idim) * sbgm -
              F(w, o->bg0_ + idim) * sbm) / rt +
             F(w, o->eb_) * F(w, o->m0_ + idim) / rt - F(w, o->b0_ + idim) * sbm / rt;
    }
    return;
  }
  if (iw > o->b0_ && iw <= o->b0_ + nc) {
    int c = iw - o->b0_;
    if (idim == c) {
      FOR_BOX(ix) { C(f) = 0.0; }
      *transport = 0;
    } else {
      FOR_BOX(ix) {
        double rt = RHOT(w);
        C(f) = -F(w, o->m0_ + c) / rt * (F(w, o->b0_ + idim) + F(w, o->bg0_ + idim)) +
               F(w, o->m0_ + idim) / rt * F(w, o->bg0_ + c);
      }
    }
    return;
  }
  /* 'Error in getflux: unknown flow variable' -- background variables are
   * never in iws for the configurations handled here */
  FOR_BOX(ix) { C(f) = 0.0; }
}

/* gradient4 (src/vacgrid.t:665-712): 5-point 4th-order derivative on ix,
 * zero on the two outermost layers of ix at every non-'mpi' boundary */
static void gradient4(sac_oracle *o, const double *q, box ix, int idim, double *gq) {
  FOR_BOX(ix) {
    C(gq) = -(SH(q, idim, 2) - 8.0 * (SH(q, idim, 1) - SH(q, idim, -1)) - SH(q, idim, -2)) /
            DX(idim) / 12.0;
  }
  /* serial build: no boundary is 'mpi', so both sides are zeroed */
  for (int k = 0; k <= 1; k++) {
    box z = ix;
    if (k == 1) {
      z.lo[idim - 1] = ix.hi[idim - 1] - 1;
    } else {
      z.hi[idim - 1] = ix.lo[idim - 1] + 1;
    }
    FOR_BOX(z) { C(gq) = 0.0; }
  }
}

/* gradient1 / gradient1L / gradient1R (src/vacusr.viscosity.t:547-675):
 * kind 0: (q(i+1)-q(i-1))/dx/2, kind 1: (q(i)-q(i-1))/dx, kind 2:
 * (q(i+1)-q(i))/dx; zero on the two outermost layers of ix unless the
 * boundary of variable 1 is 'periodic' (or 'mpi') */
static void gradient1k(sac_oracle *o, const double *q, box ix, int idim, int kind, double *gq) {
  FOR_BOX(ix) {
    if (kind == 0)
      C(gq) = (SH(q, idim, 1) - SH(q, idim, -1)) / DX(idim) / 2.0;
    else if (kind == 1)
      C(gq) = (C(q) - SH(q, idim, -1)) / DX(idim);
    else
      C(gq) = (SH(q, idim, 1) - C(q)) / DX(idim);
  }
  for (int k = 0; k <= 1; k++) {
    int iB = 2 * (idim - 1) + k;
    if (o->typeB[0][iB] == OR_BC_PERIODIC) continue;
    box z = ix;
    if (k == 1) {
      z.lo[idim - 1] = ix.hi[idim - 1] - 1;
    } else {
      z.hi[idim - 1] = ix.lo[idim - 1] + 1;
    }
    FOR_BOX(z) { C(gq) = 0.0; }
  }
}

/* ------------------------------------------------------------------ */
/* user module sources                                                  */

/* addsource_grav (src/vacusr.gravity.t:24-81): gravity on the perturbed
 * density only */
static void addsource_grav(sac_oracle *o, double qdt, box ixO, const double *w, double *wnew) {
  for (int ii = 0; ii < o->niws; ii++) {
    int iw = o->iws[ii];
    if (iw > o->m0_ && iw <= o->m0_ + o->ndim) {
      int idim = iw - o->m0_;
      double g = eq_grav(o, idim);
      if (fabs(g) > SMALLDOUBLE) {
        FOR_BOX(ixO) { F(wnew, o->m0_ + idim) = F(wnew, o->m0_ + idim) + qdt * g * F(w, o->rho_); }
      }
    } else if (iw == o->e_) {
      for (int idim = 1; idim <= o->ndim; idim++) {
        double g = eq_grav(o, idim);
        if (fabs(g) > SMALLDOUBLE) {
          FOR_BOX(ixO) {
            F(wnew, o->e_) = F(wnew, o->e_) + qdt * g * F(w, o->rho_) * F(w, o->m0_ + idim) / RHOT(w);
          }
        }
      }
    }
  }
}

/* setnu (src/vacusr.viscosity.t:224-455): hyperviscosity coefficients
 * nuR/nuL on ix (= the mesh) for variable iw (173 = thermal energy) */
static void setnu(sac_oracle *o, const double *w, int iw, int idim, box ix, double *nuR,
                  double *nuL) {
  int new_cmax = 1;
  getcmax(o, &new_cmax, w, ix, idim, o->sn_cmax);
  double c_tot = -BIGDOUBLE;
  FOR_BOX(ix) {
    if (C(o->sn_cmax) > c_tot) c_tot = C(o->sn_cmax);
  }
  double c_hyp = 0.4;
  if (iw > o->b0_ && iw <= o->b0_ + o->ndim) c_hyp = 0.04;
  if (iw == o->rho_) c_hyp = 0.04;
  if (iw == 173) c_hyp = 0.04;

  box g = grid_box(o);
  double *tn = o->sn_tmp_nu;
  if (iw != 173) {
    FOR_BOX(g) { C(tn) = F(w, iw); }
    if (iw > o->m0_ && iw <= o->m0_ + o->ndim) FOR_BOX(g) { C(tn) = F(w, iw) / RHOT(w); }
  }
  if (iw == 173) {
    FOR_BOX(g) {
      double sb = 0.0, sm = 0.0;
      for (int c = 1; c <= o->ndir; c++) {
        double b = F(w, o->b0_ + c), m = F(w, o->m0_ + c);
        sb = (c == 1) ? b * b : sb + b * b;
        sm = (c == 1) ? m * m : sm + m * m;
      }
      C(tn) = F(w, o->e_) - 0.5 * (sb + sm / RHOT(w));
    }
  }
  /* tmp_nuI(ixF^T)=tmp_nu(ixY^T), ixY=ix^LADD2, ixF=ixY+1 (index shift +1) */
  double *tI = o->sn_tmp_nuI;
  int nd = o->ndim;
  box ixY = box_add(o, ix, 2);
  box ixFLL = ixY;
  for (int d = 0; d < nd; d++) {
    ixFLL.lo[d] += 1;
    ixFLL.hi[d] += 1;
  }
  FOR_BOX(ixY) { tI[IDXE(o, i1 + (nd > 0), i2 + (nd > 1), i3 + (nd > 2))] = C(tn); }
  int iwc = (iw == 173) ? o->e_ : iw;
  for (int k = 0; k <= 1; k++) {
    int iB = 2 * (idim - 1) + k;
    (void)iwc; /* serial: typeB(iwc,iB) is never 'mpi' */
    box s = ixFLL;
    int d = idim - 1;
    if (k == 1) {
      s.lo[d] = s.hi[d] = ixFLL.hi[d] + 1;
      FOR_BOX(s) {
        int j1 = i1, j2 = i2, j3 = i3;
        if (d == 0) j1 = ixFLL.hi[0] - 5;
        if (d == 1) j2 = ixFLL.hi[1] - 5;
        if (d == 2) j3 = ixFLL.hi[2] - 5;
        tI[IDXE(o, i1, i2, i3)] = tI[IDXE(o, j1, j2, j3)];
      }
    } else {
      s.lo[d] = s.hi[d] = ixFLL.lo[d] - 1;
      FOR_BOX(s) {
        int j1 = i1, j2 = i2, j3 = i3;
        if (d == 0) j1 = ixFLL.lo[0] + 5;
        if (d == 1) j2 = ixFLL.lo[1] + 5;
        if (d == 2) j3 = ixFLL.lo[2] + 5;
        tI[IDXE(o, i1, i2, i3)] = tI[IDXE(o, j1, j2, j3)];
      }
    }
    (void)iB;
  }
  box ixF = ixFLL; /* ixF^L=ixF^LL^LSUB1 */
  for (int d = 0; d < nd; d++) {
    ixF.lo[d] += 1;
    ixF.hi[d] -= 1;
  }
  int e1 = (idim == 1), e2 = (idim == 2), e3 = (idim == 3);
#define TI(s) tI[IDXE(o, i1 + e1 * (s), i2 + e2 * (s), i3 + e3 * (s))]
  FOR_BOX(ixF) {
    o->sn_d3R[IDXD(o, i1, i2, i3)] = fabs(3.0 * (TI(1) - TI(0)) - (TI(2) - TI(-1)));
    o->sn_d3L[IDXD(o, i1, i2, i3)] = fabs(3.0 * (TI(0) - TI(-1)) - (TI(1) - TI(-2)));
  }
  FOR_BOX(ixFLL) {
    o->sn_d1R[IDXD(o, i1, i2, i3)] = fabs(TI(1) - TI(0));
    o->sn_d1L[IDXD(o, i1, i2, i3)] = fabs(TI(0) - TI(-1));
  }
#undef TI
  int s1 = (nd > 0), s2 = (nd > 1), s3 = (nd > 2);
  FOR_BOX(ix) {
    double m3r = -BIGDOUBLE, m1r = -BIGDOUBLE, m3l = -BIGDOUBLE, m1l = -BIGDOUBLE;
    for (int s = -1; s <= 1; s++) {
      long id = IDXD(o, i1 + s1 + e1 * s, i2 + s2 + e2 * s, i3 + s3 + e3 * s);
      if (o->sn_d3R[id] > m3r) m3r = o->sn_d3R[id];
      if (o->sn_d3L[id] > m3l) m3l = o->sn_d3L[id];
    }
    for (int s = -2; s <= 2; s++) {
      long id = IDXD(o, i1 + s1 + e1 * s, i2 + s2 + e2 * s, i3 + s3 + e3 * s);
      if (o->sn_d1R[id] > m1r) m1r = o->sn_d1R[id];
      if (o->sn_d1L[id] > m1l) m1l = o->sn_d1L[id];
    }
    C(o->sn_md3R) = m3r;
    C(o->sn_md1R) = m1r;
    C(o->sn_md3L) = m3l;
    C(o->sn_md1L) = m1l;
  }
  double mx = -BIGDOUBLE;
  FOR_BOX(ix) {
    if (C(o->sn_md1R) > 0.0)
      C(nuR) = c_tot * c_hyp * C(o->sn_md3R) / C(o->sn_md1R) * DX(idim);
    else
      C(nuR) = 0.0;
    if (C(nuR) > mx) mx = C(nuR);
  }
  if (mx > o->maxviscoef) o->maxviscoef = mx;
  mx = -BIGDOUBLE;
  FOR_BOX(ix) {
    if (C(o->sn_md1L) > 0.0)
      C(nuL) = c_tot * c_hyp * C(o->sn_md3L) / C(o->sn_md1L) * DX(idim);
    else
      C(nuL) = 0.0;
    if (C(nuL) > mx) mx = C(nuL);
  }
  if (mx > o->maxviscoef) o->maxviscoef = mx;
}

/* addsource_visc (src/vacusr.viscosity.t:2-221): hyperdiffusion sources
 * from w (= wCT) added to wnew over the whole grid ixI */
static void addsource_visc(sac_oracle *o, double qdt, box ixI, box ixO, const double *w,
                           double *wnew) {
  int nd = o->ndim;
  /* ensurebound(2,ixI,ixO): ixI=ixG is wider than ixO=ixM by 2 -> no-op */
  box ix = box_add(o, ixO, 1);
  /* setnushk: nushk(ix,idim)=0 (body commented out in the reference) */
  for (int idim = 1; idim <= nd; idim++) {
    double *nk = o->av_nushk + (long)(idim - 1) * o->n;
    FOR_BOX(ix) { C(nk) = 0.0; }
  }
  double *tmp = o->tmp, *tmp2 = o->tmp2;
  double *nuL = o->av_nuL, *nuR = o->av_nuR, *tL = o->av_L, *tR = o->av_R, *tC = o->av_C;
#define NUSHK(idim) (o->av_nushk[(long)((idim)-1) * o->n + IDX(o, i1, i2, i3)])
  for (int idim = 1; idim <= nd; idim++) {
    FOR_BOX(ixI) { C(tmp) = F(w, o->rho_); }
    setnu(o, w, o->rho_, idim, ixO, nuR, nuL);
    gradient1k(o, tmp, ix, idim, 1, tmp2);
    FOR_BOX(ixI) { C(tL) = (C(nuL) + NUSHK(idim)) * C(tmp2); }
    gradient1k(o, tmp, ix, idim, 2, tmp2);
    FOR_BOX(ixI) { C(tR) = (C(nuR) + NUSHK(idim)) * C(tmp2); }
    FOR_BOX(ixI) { F(wnew, o->rho_) = F(wnew, o->rho_) + (C(tR) - C(tL)) / DX(idim) * qdt; }
  }
  for (int idim = 1; idim <= nd; idim++) {
    FOR_BOX(ixI) {
      double sb = 0.0, sm = 0.0;
      for (int c = 1; c <= o->ndir; c++) {
        double b = F(w, o->b0_ + c), m = F(w, o->m0_ + c);
        sb = (c == 1) ? b * b : sb + b * b;
        sm = (c == 1) ? m * m : sm + m * m;
      }
      C(tmp) = F(w, o->e_) - 0.5 * (sb + sm / RHOT(w));
    }
    setnu(o, w, 173, idim, ixO, nuR, nuL);
    gradient1k(o, tmp, ix, idim, 1, tmp2);
    FOR_BOX(ixI) { C(tL) = (C(nuL) + NUSHK(idim)) * C(tmp2); }
    gradient1k(o, tmp, ix, idim, 2, tmp2);
    FOR_BOX(ixI) { C(tR) = (C(nuR) + NUSHK(idim)) * C(tmp2); }
    FOR_BOX(ixI) { F(wnew, o->e_) = F(wnew, o->e_) + (C(tR) - C(tL)) / DX(idim) * qdt; }
  }
  FOR_BOX(ixI) { C(o->av_rhoC) = F(w, o->rho_) + F(w, o->rhob_); }
  for (int k = 1; k <= nd; k++) {
    FOR_BOX(ix) {
      C(o->av_rhoL) = ((F(w, o->rho_) + F(w, o->rhob_)) + (FSH(w, o->rho_, k, -1) + FSH(w, o->rhob_, k, -1))) / 2.0;
      C(o->av_rhoR) = ((FSH(w, o->rho_, k, 1) + FSH(w, o->rhob_, k, 1)) + (F(w, o->rho_) + F(w, o->rhob_))) / 2.0;
    }
    for (int l = 1; l <= nd; l++) {
      setnu(o, w, l + o->m0_, k, ixO, nuR, nuL);
      FOR_BOX(ixI) { C(tmp) = F(w, o->m0_ + l) / RHOT(w); }
      for (int ii1 = 0; ii1 <= 1; ii1++) {
        int i, ii0;
        if (ii1 == 0) {
          i = k;
          ii0 = l;
        } else {
          i = l;
          ii0 = k;
        }
        if (i == k) {
          FOR_BOX(ix) {
            C(o->av_VL) = (F(w, o->m0_ + ii0) + FSH(w, o->m0_ + ii0, k, -1)) / 2.0;
            C(o->av_VR) = (FSH(w, o->m0_ + ii0, k, 1) + F(w, o->m0_ + ii0)) / 2.0;
          }
          gradient1k(o, tmp, ix, k, 1, tmp2);
          FOR_BOX(ixI) { C(tL) = (C(nuL) + NUSHK(k)) * C(tmp2); }
          gradient1k(o, tmp, ix, k, 2, tmp2);
          FOR_BOX(ixI) { C(tR) = (C(nuR) + NUSHK(k)) * C(tmp2); }
          FOR_BOX(ixI) { C(tmp2) = (C(o->av_rhoR) * C(tR) - C(o->av_rhoL) * C(tL)) / DX(k) / 2.0; }
          FOR_BOX(ixI) { F(wnew, o->m0_ + ii0) = F(wnew, o->m0_ + ii0) + C(tmp2) * qdt; }
          FOR_BOX(ixI) { C(tmp2) = (C(o->av_VR) * C(tR) - C(o->av_VL) * C(tL)) / DX(k) / 2.0; }
          FOR_BOX(ixI) { F(wnew, o->e_) = F(wnew, o->e_) + C(tmp2) * qdt; }
        }
        if (i != k) {
          gradient1k(o, tmp, ix, k, 0, tmp2);
          FOR_BOX(ixI) { C(tmp2) = C(tmp2) * (C(nuL) + C(nuR) + 2.0 * NUSHK(k)) / 2.0 / 2.0; }
          FOR_BOX(ixI) { C(tmp) = C(o->av_rhoC) * C(tmp2); }
          gradient1k(o, tmp, ix, i, 0, tC);
          FOR_BOX(ixI) { F(wnew, o->m0_ + ii0) = F(wnew, o->m0_ + ii0) + C(tC) * qdt; }
          FOR_BOX(ixI) { C(tmp) = F(w, o->m0_ + ii0) * C(tmp2); }
          gradient1k(o, tmp, ix, i, 0, tC);
          FOR_BOX(ixI) { F(wnew, o->e_) = F(wnew, o->e_) + C(tC) * qdt; }
        }
      }
    }
  }
  for (int k = 1; k <= nd; k++) {
    for (int l = 1; l <= nd; l++) {
      if (k == l) continue;
      setnu(o, w, o->b0_ + l, k, ixO, nuR, nuL);
      for (int ii1 = 0; ii1 <= 1; ii1++) {
        int ii0, m, j;
        double sB;
        if (ii1 == 0) {
          ii0 = k;
          m = l;
          sB = -1.0;
          j = k;
        } else {
          ii0 = l;
          m = k;
          sB = 1.0;
          j = l;
        }
        if (m == k) {
          FOR_BOX(ix) {
            C(o->av_BL) = (F(w, o->b0_ + j) + FSH(w, o->b0_ + j, m, -1)) / 2.0;
            C(o->av_BR) = (FSH(w, o->b0_ + j, m, 1) + F(w, o->b0_ + j)) / 2.0;
          }
          FOR_BOX(ixI) { C(tmp) = F(w, o->b0_ + l); }
          gradient1k(o, tmp, ix, k, 1, tmp2);
          FOR_BOX(ixI) { C(tL) = C(nuL) * C(tmp2); }
          gradient1k(o, tmp, ix, k, 2, tmp2);
          FOR_BOX(ixI) { C(tR) = C(nuR) * C(tmp2); }
          FOR_BOX(ixI) {
            F(wnew, o->b0_ + ii0) = F(wnew, o->b0_ + ii0) + sB * (C(tR) - C(tL)) / DX(k) * qdt;
          }
          FOR_BOX(ixI) {
            F(wnew, o->e_) = F(wnew, o->e_) +
                             sB * (C(tR) * C(o->av_BR) - C(tL) * C(o->av_BL)) / DX(k) * qdt;
          }
        }
        if (m != k) {
          FOR_BOX(ixI) { C(tmp) = F(w, o->b0_ + l); }
          gradient1k(o, tmp, ix, k, 0, tmp2);
          FOR_BOX(ixI) { C(tmp2) = C(tmp2) * (C(nuL) + C(nuR)) / 2.0; }
          gradient1k(o, tmp2, ix, m, 0, tC);
          FOR_BOX(ixI) { F(wnew, o->b0_ + ii0) = F(wnew, o->b0_ + ii0) + sB * C(tC) * qdt; }
          FOR_BOX(ixI) { C(tmp2) = C(tmp2) * F(w, o->b0_ + j); }
          gradient1k(o, tmp2, ix, m, 0, tC);
          FOR_BOX(ixI) { F(wnew, o->e_) = F(wnew, o->e_) + sB * C(tC) * qdt; }
        }
      }
    }
  }
#undef NUSHK
}

/* specialsource: gravity, then hyperdiffusion if |eqpar(nu_)|>smalldouble
 * (user module order of src/vacusr.f:1279-1287 without the Slog driver) */
static void specialsource(sac_oracle *o, double qdt, box ixI, box ixO, const double *wCT,
                          double *w) {
  addsource_grav(o, qdt, ixO, wCT, w);
  if (fabs(eq_nu(o)) > SMALLDOUBLE) addsource_visc(o, qdt, ixI, ixO, wCT, w);
}

/* getdivb (src/vacphys.mhd0.t:106-146) with fourthorder=F: tmp = b+bg on
 * ixO+1 and gradient4 along each axis (which zeroes its two outermost
 * cells, the only ones that would read tmp beyond ixO+1) */
static void getdivb(sac_oracle *o, const double *w, box ixO, double *divb) {
  box ix = box_add(o, ixO, 1);
  FOR_BOX(ixO) { C(divb) = 0.0; }
  for (int idim = 1; idim <= o->ndim; idim++) {
    FOR_BOX(ix) { C(o->tmp) = F(w, o->b0_ + idim) + F(w, o->bg0_ + idim); }
    gradient4(o, o->tmp, ixO, idim, o->tmp2);
    FOR_BOX(ixO) { C(divb) = C(divb) + C(o->tmp2); }
  }
}

/* addsource_divb (src/vacphys.mhd0.t:264-297): Powell's div B sources,
 * m -= (b+bg) qdt divB, b -= m/rho_t qdt divB, e -= (m.(b+bg))/rho_t qdt divB */
static void addsource_divb(sac_oracle *o, double qdt, box ixO, const double *w, double *wnew) {
  double *divb = o->db_divb;
  getdivb(o, w, ixO, divb);
  FOR_BOX(ixO) { C(divb) = qdt * C(divb); }
  for (int ii = 0; ii < o->niws; ii++) {
    int iw = o->iws[ii];
    if (iw > o->m0_ && iw <= o->m0_ + o->ndir) {
      int c = iw - o->m0_;
      FOR_BOX(ixO) { F(wnew, iw) = F(wnew, iw) - (F(w, o->b0_ + c) + F(w, o->bg0_ + c)) * C(divb); }
    } else if (iw > o->b0_ && iw <= o->b0_ + o->ndir) {
      int c = iw - o->b0_;
      FOR_BOX(ixO) {
        F(wnew, iw) = F(wnew, iw) - F(w, o->m0_ + c) / (F(w, o->rho_) + F(w, o->rhob_)) * C(divb);
      }
    } else if (iw == o->e_) {
      FOR_BOX(ixO) {
        double s = 0.0;
        for (int c = 1; c <= o->ndir; c++) {
          double t = F(w, o->m0_ + c) * (F(w, o->b0_ + c) + F(w, o->bg0_ + c));
          s = (c == 1) ? t : s + t;
        }
        F(wnew, iw) = F(wnew, iw) - s / (F(w, o->rho_) + F(w, o->rhob_)) * C(divb);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* resistive MHD with a constant eta = eqpar(eta_) > 0
 * (src/vacphys.mhdres.t; eta < 0 needs the user's specialeta) */

static double eq_eta(const sac_oracle *o) { return o->eqpar[1]; }

/* Levi-Civita symbol lvc(i,j,k), 1-based */
static int lvc(int i, int j, int k) {
  if (i == j || j == k || i == k) return 0;
  return ((i == 1 && j == 2) || (i == 2 && j == 3) || (i == 3 && j == 1)) ? 1 : -1;
}

/* the module's own gradient (src/vacphys.mhdres.t:304-317): the NEGATIVE
 * central difference -(q(i+1)-q(i-1))/dx/2, no edge zeroing */
static void gradient_res(sac_oracle *o, const double *q, box ix, int idim, double *gq) {
  FOR_BOX(ix) { C(gq) = -(SH(q, idim, 1) - SH(q, idim, -1)) / DX(idim) / 2.0; }
}

/* getcurrent (src/vacphys.mhdres.t:48-93): J_i = sum lvc(i,j,k) gradient(b_k+bg_k)_j
 * on ix, components idirmin0..3 (idirmin0 = 7-2 ndir); returns idirmin */
static int getcurrent(sac_oracle *o, const double *w, box ix) {
  int idirmin0 = 7 - 2 * o->ndir, idirmin = 4;
  box ixI = box_add(o, ix, 1);
  for (int idir = idirmin0; idir <= 3; idir++) {
    double *J = o->rs_current + (long)(idir - 1) * o->n;
    FOR_BOX(ix) { C(J) = 0.0; }
  }
  for (int idir = idirmin0; idir <= 3; idir++)
    for (int jdir = 1; jdir <= o->ndim; jdir++)
      for (int kdir = 1; kdir <= o->ndir; kdir++) {
        int l = lvc(idir, jdir, kdir);
        if (l == 0) continue;
        double *J = o->rs_current + (long)(idir - 1) * o->n;
        FOR_BOX(ixI) { C(o->tmp) = F(w, o->b0_ + kdir) + F(w, o->bg0_ + kdir); }
        gradient_res(o, o->tmp, ix, jdir, o->tmp2);
        if (l == 1) {
          FOR_BOX(ix) { C(J) = C(J) + C(o->tmp2); }
        } else {
          FOR_BOX(ix) { C(J) = C(J) - C(o->tmp2); }
        }
        if (idir < idirmin) idirmin = idir;
      }
  return idirmin;
}

/* addsource_res2 (src/vacphys.mhdres.t:223-301), conservative 5-point form:
 * b_i -= sum lvc(i,j,k) gradient(J_k eta qdt)_j,
 * e += sum lvc(i,j,k) gradient((b_j+bg_j) J_k eta qdt)_i */
static void addsource_res2(sac_oracle *o, double qdt, box ixO, const double *w, double *wnew) {
  box ix = box_add(o, ixO, 1);
  int idirmin = getcurrent(o, w, ix);
  double eta = eq_eta(o);
  for (int ii = 0; ii < o->niws; ii++) {
    int iw = o->iws[ii];
    if (iw > o->b0_ && iw <= o->b0_ + o->ndir) {
      int idir = iw - o->b0_;
      for (int jdir = 1; jdir <= o->ndim; jdir++)
        for (int kdir = idirmin; kdir <= 3; kdir++) {
          int l = lvc(idir, jdir, kdir);
          if (l == 0) continue;
          const double *J = o->rs_current + (long)(kdir - 1) * o->n;
          FOR_BOX(ix) { C(o->tmp) = C(J) * eta * qdt; }
          gradient_res(o, o->tmp, ixO, jdir, o->tmp2);
          if (l == 1) {
            FOR_BOX(ixO) { F(wnew, iw) = F(wnew, iw) - C(o->tmp2); }
          } else {
            FOR_BOX(ixO) { F(wnew, iw) = F(wnew, iw) + C(o->tmp2); }
          }
        }
    } else if (iw == o->e_) {
      for (int idir = 1; idir <= o->ndim; idir++)
        for (int jdir = 1; jdir <= o->ndir; jdir++)
          for (int kdir = idirmin; kdir <= 3; kdir++) {
            int l = lvc(idir, jdir, kdir);
            if (l == 0) continue;
            const double *J = o->rs_current + (long)(kdir - 1) * o->n;
            FOR_BOX(ix) {
              C(o->tmp) = (F(w, o->b0_ + jdir) + F(w, o->bg0_ + jdir)) * C(J) * eta * qdt;
            }
            gradient_res(o, o->tmp, ixO, idir, o->tmp2);
            if (l == 1) {
              FOR_BOX(ixO) { F(wnew, iw) = F(wnew, iw) + C(o->tmp2); }
            } else {
              FOR_BOX(ixO) { F(wnew, iw) = F(wnew, iw) - C(o->tmp2); }
            }
          }
    }
  }
}

/* addsource_res1 (src/vacphys.mhdres.t:96-220), compact 3-point form:
 * b_i += qdt eta lapl(b_i+bg_i), e += qdt eta lapl(b_i+bg_i) (b_i+bg_i),
 * then e += qdt eta J^2 */
static void addsource_res1(sac_oracle *o, double qdt, box ix, const double *w, double *wnew) {
  int idirmin = getcurrent(o, w, ix);
  double eta = eq_eta(o);
  box g = grid_box(o);
  for (int idir = 1; idir <= o->ndir; idir++) {
    FOR_BOX(ix) { C(o->tmp) = 0.0; }
    FOR_BOX(g) { C(o->tmp2) = F(w, o->b0_ + idir) + F(w, o->bg0_ + idir); }
    for (int idim = 1; idim <= o->ndim; idim++) {
      FOR_BOX(ix) {
        C(o->tmp) = C(o->tmp) + (SH(o->tmp2, idim, 1) - 2 * C(o->tmp2) + SH(o->tmp2, idim, -1)) /
                                    (DX(idim) * DX(idim));
      }
    }
    FOR_BOX(ix) { C(o->tmp) = C(o->tmp) * eta; }
    for (int ii = 0; ii < o->niws; ii++) {
      int iw = o->iws[ii];
      if (iw == o->b0_ + idir) {
        FOR_BOX(ix) { F(wnew, iw) = F(wnew, iw) + qdt * C(o->tmp); }
      } else if (iw == o->e_) {
        FOR_BOX(ix) {
          F(wnew, iw) = F(wnew, iw) + qdt * C(o->tmp) * (F(w, o->b0_ + idir) + F(w, o->bg0_ + idir));
        }
      }
    }
  }
  for (int ii = 0; ii < o->niws; ii++) {
    if (o->iws[ii] != o->e_) continue;
    FOR_BOX(ix) { C(o->tmp) = 0.0; }
    for (int idir = idirmin; idir <= 3; idir++) {
      const double *J = o->rs_current + (long)(idir - 1) * o->n;
      FOR_BOX(ix) { C(o->tmp) = C(o->tmp) + C(J) * C(J); }
    }
    FOR_BOX(ix) { F(wnew, o->e_) = F(wnew, o->e_) + qdt * eta * C(o->tmp); }
  }
}

/* getdt_res (src/vacphys.mhdres.t:5-44) for eta > 0: dtdiffpar min(dx)^2/eta */
static double getdt_res(sac_oracle *o, double dt) {
  double eta = eq_eta(o);
  if (eta == 0.0) return dt;
  box ix = mesh_box(o);
  double mn = BIGDOUBLE;
  for (int idim = 1; idim <= o->ndim; idim++) FOR_BOX(ix) {
      if (DX(idim) < mn) mn = DX(idim);
    }
  return fmin(dt, o->dtdiffpar * mn * mn / eta);
}

/* addsource2 (src/vac.t:552-582): specialsource, then addsource
 * (src/vacphys.mhd0.t:225-260): the resistive sources when eta /= 0, then
 * Powell's sources when divbfix and ndim > 1 */
static void addsource2(sac_oracle *o, double qdt, box ixI, box ixO, const double *wCT,
                       double *w) {
  specialsource(o, qdt, ixI, ixO, wCT, w);
  if (fabs(eq_eta(o)) > SMALLDOUBLE) {
    if (o->compactres)
      addsource_res1(o, qdt, ixO, wCT, w);
    else
      addsource_res2(o, qdt, ixO, wCT, w);
  }
  if (o->divbfix && o->ndim > 1) addsource_divb(o, qdt, ixO, wCT, w);
}

void oracle_set_resistive(sac_oracle *o, int compactres, double dtdiffpar) {
  o->compactres = compactres ? 1 : 0;
  o->dtdiffpar = dtdiffpar;
}

void oracle_set_divbfix(sac_oracle *o, int on) { o->divbfix = on ? 1 : 0; }

/* centdiff4 (src/vaccd.t:7-78) with idim^LIM = 1..ndim */
static void centdiff4(sac_oracle *o, double qdt, box ixI, box ixO, const double *wCT,
                      double *w) {
  int nd = o->ndim;
  for (int idim = 1; idim <= nd; idim++) {
    box ix = box_add_dir(ixO, idim, 2);
    getv(o, wCT, ix, idim, o->cd_v);
    for (int ii = 0; ii < o->niws; ii++) {
      int iw = o->iws[ii];
      int transport;
      getflux(o, wCT, ix, iw, idim, o->cd_f, &transport);
      if (transport) FOR_BOX(ix) { C(o->cd_f) = C(o->cd_f) + C(o->cd_v) * F(wCT, iw); }
      gradient4(o, o->cd_f, ixO, idim, o->tmp);
      FOR_BOX(ix) { F(w, iw) = F(w, iw) - qdt * C(o->tmp); }
      if (iw == o->e_) {
        gradient4(o, o->cd_v, ixO, idim, o->tmp);
        getptotal_bg(o, w, ix, o->cd_fb);
        FOR_BOX(ix) { F(w, iw) = F(w, iw) - qdt * C(o->tmp) * C(o->cd_fb); }
        for (int idir = 1; idir <= nd; idir++) {
          gradient4(o, o->cd_v, ixO, idir, o->tmp);
          FOR_BOX(ix) {
            F(w, iw) = F(w, iw) + qdt * F(w, o->bg0_ + idir) * F(w, o->bg0_ + idim) * C(o->tmp);
          }
        }
      }
    }
  }
  /* sourceunsplit: addsource2(qdt*(idimmax-idimmin+one)/ndim, ...) */
  double q2 = qdt * (double)(nd) / (double)nd;
  addsource2(o, q2, ixI, ixO, wCT, w);
}

/* advect1 with method 'cd4' (src/vac.t:490-549) */
static void advect1(sac_oracle *o, double qdt, const double *wCT, double *w) {
  box ixI = grid_box(o);
  box ixO = box_add(o, ixI, -2);
  centdiff4(o, qdt, ixI, ixO, wCT, w);
  getboundary(o, w);
}

/* advect (src/vac.t:364-487): the Runge-Kutta family on the full grid */
int oracle_advance(sac_oracle *o, double dt, int typeadvance) {
  long nn = (long)o->nw * o->n;
  double *w = o->w, *w1 = o->w1, *w2 = o->w2, *w3 = o->w3;
  memcpy(w1, w, sizeof(double) * (size_t)nn);
  switch (typeadvance) {
    case OR_ADV_ONESTEP:
      advect1(o, dt, w1, w);
      break;
    case OR_ADV_TWOSTEP:
      /* with typepred1='nul' the reference reduces twostep to onestep
       * (src/vacio.t:502); a 'cd4' predictor is restated literally */
      advect1(o, dt / 2, w, w1);
      advect1(o, dt, w1, w);
      break;
    case OR_ADV_THREESTEP:
      advect1(o, dt, w, w1);
      for (long i = 0; i < nn; i++) w2[i] = 3 * 0.25 * w[i] + 0.25 * w1[i];
      advect1(o, dt / 4, w1, w2);
      for (long i = 0; i < nn; i++) w[i] = w[i] / 3 + w2[i] * (2.0 / 3);
      advect1(o, dt * 2.0 / 3, w2, w);
      break;
    case OR_ADV_FOURSTEP:
      advect1(o, dt / 2, w, w1);
      memcpy(w2, w, sizeof(double) * (size_t)nn);
      advect1(o, dt / 2, w1, w2);
      memcpy(w3, w, sizeof(double) * (size_t)nn);
      advect1(o, dt, w2, w3);
      for (long i = 0; i < nn; i++) w1[i] = (w1[i] + 2 * w2[i] + w3[i] - 4 * w[i]) / 3;
      advect1(o, dt / 6, w3, w);
      for (long i = 0; i < nn; i++) w[i] = w[i] + w1[i];
      break;
    case OR_ADV_STERCK:
      advect1(o, dt * (double)0.12f, w, w1); /* 0.12 is a default-real literal */
      memcpy(w2, w, sizeof(double) * (size_t)nn);
      advect1(o, dt / 4, w1, w2);
      memcpy(w1, w, sizeof(double) * (size_t)nn);
      advect1(o, dt / 2, w2, w1);
      advect1(o, dt, w1, w);
      break;
    case OR_ADV_JAMESON:
      advect1(o, dt / 4, w, w1);
      memcpy(w2, w, sizeof(double) * (size_t)nn);
      advect1(o, dt / 3, w1, w2);
      memcpy(w1, w, sizeof(double) * (size_t)nn);
      advect1(o, dt / 2, w2, w1);
      advect1(o, dt, w1, w);
      break;
    default:
      return -1;
  }
  return 0;
}

void oracle_advect1_on(sac_oracle *o, double qdt, const double *wCT, double *w) {
  advect1(o, qdt, wCT, w);
}

/* ------------------------------------------------------------------ */
/* time step                                                            */

/* getdt_courant (src/vac.t:641-699) */
double oracle_getdt_courant(sac_oracle *o, double courantpar, double *courantmax_out) {
  box ix = mesh_box(o);
  double dt = BIGDOUBLE;
  double courantmax = 0.0;
  int new_cmax = 1;
  for (int idim = 1; idim <= o->ndim; idim++) {
    getcmax(o, &new_cmax, o->w, ix, idim, o->gc_cmax);
    double mx = -BIGDOUBLE;
    FOR_BOX(ix) {
      double v = C(o->gc_cmax) / DX(idim);
      if (v > mx) mx = v;
    }
    if (mx > courantmax) courantmax = mx;
  }
  if (courantmax > SMALLDOUBLE) dt = fmin(dt, courantpar / courantmax);
  if (courantmax_out) *courantmax_out = courantmax;
  return dt;
}

/* getdt_visc (src/vacusr.viscosity.t:509-541): uses maxviscoef collected
 * since the previous call, then resets it */
static double getdt_visc(sac_oracle *o, double dt) {
  box ix = mesh_box(o);
  for (int idim = 1; idim <= o->ndim; idim++) {
    double *nk = o->dv_nushk + (long)(idim - 1) * o->n;
    FOR_BOX(ix) { C(nk) = 0.0; }
  }
  /* dtdiffpar=0.25d0 assigns the common variable, which the resistive dt
   * limit of later steps then uses */
  o->dtdiffpar = 0.25;
  double dtdiffpar = o->dtdiffpar;
  for (int idim = 1; idim <= o->ndim; idim++) {
    double *nk = o->dv_nushk + (long)(idim - 1) * o->n;
    FOR_BOX(ix) { C(o->dv_tmpdt) = o->maxviscoef + C(nk); }
    double mx = -BIGDOUBLE;
    FOR_BOX(ix) {
      double v = C(o->dv_tmpdt) / (DX(idim) * DX(idim));
      if (v > mx) mx = v;
    }
    double dtd = dtdiffpar / mx; /* +Inf when maxviscoef == 0 */
    dt = fmin(dt, dtd);
  }
  o->maxviscoef = 0.0;
  return dt;
}

/* getdt_grav (src/vacusr.gravity.t:97-138): computed at it==itmin, saved */
static double getdt_grav(sac_oracle *o, double dt) {
  box ix = mesh_box(o);
  if (o->it == o->itmin) {
    o->dtgrav = BIGDOUBLE;
    for (int idim = 1; idim <= o->ndim; idim++) {
      double g = eq_grav(o, idim);
      if (fabs(g) > 0.0) {
        double mx = -BIGDOUBLE;
        for (int jd = 1; jd <= o->ndim; jd++) FOR_BOX(ix) {
            double v = fabs(g) / DX(jd);
            if (v > mx) mx = v;
          }
        double c = 1.0 / sqrt(mx);
        if (c < o->dtgrav) o->dtgrav = c;
      }
    }
  }
  return fmin(dt, o->dtgrav);
}

double oracle_getdt_res(sac_oracle *o, double dt) { return getdt_res(o, dt); }

double oracle_getdt_special(sac_oracle *o, double dt) {
  if (fabs(eq_nu(o)) > SMALLDOUBLE) dt = getdt_visc(o, dt);
  dt = getdt_grav(o, dt);
  return dt;
}

double oracle_maxviscoef(const sac_oracle *o) { return o->maxviscoef; }
void oracle_set_maxviscoef(sac_oracle *o, double v) { o->maxviscoef = v; }

void oracle_set_time(sac_oracle *o, int it, int itmin, double t) {
  o->it = it;
  o->itmin = itmin;
  o->t = t;
}
double oracle_time(const sac_oracle *o) { return o->t; }
int oracle_it(const sac_oracle *o) { return o->it; }

/* one iteration of the main loop (src/vac.t:60-104) with dtpar<0,
 * dtcantgrow=F, slowsteps=0, tmaxexact=F */
double oracle_step(sac_oracle *o, double courantpar, int typeadvance) {
  double dt = BIGDOUBLE;
  if (courantpar > 0.0) dt = oracle_getdt_courant(o, courantpar, NULL);
  /* getdt (src/vacphys.mhd0.t:83-103): the resistive limit when eta /= 0 */
  dt = getdt_res(o, dt);
  dt = oracle_getdt_special(o, dt);
  oracle_advance(o, dt, typeadvance);
  o->it += 1;
  o->t += dt;
  return dt;
}

/* ------------------------------------------------------------------ */
/* building blocks exposed for tests                                    */

void oracle_gradient4(sac_oracle *o, const double *q, int idim, double *gq) {
  gradient4(o, q, mesh_box(o), idim, gq);
}

void oracle_gradient1(sac_oracle *o, const double *q, int idim, int kind, double *gq) {
  gradient1k(o, q, box_add(o, mesh_box(o), 1), idim, kind, gq);
}

void oracle_getflux(sac_oracle *o, const double *w, int iw, int idim, double *f,
                    int *transport) {
  getflux(o, w, box_add_dir(mesh_box(o), idim, 2), iw, idim, f, transport);
}

void oracle_getcmax(sac_oracle *o, const double *w, int idim, double *cmax) {
  int nc = 1;
  getcmax(o, &nc, w, mesh_box(o), idim, cmax);
}

void oracle_setnu(sac_oracle *o, const double *w, int iw, int idim, double *nuR,
                  double *nuL) {
  setnu(o, w, iw, idim, mesh_box(o), nuR, nuL);
}

void oracle_addsource_visc(sac_oracle *o, double qdt, const double *w, double *wnew) {
  addsource_visc(o, qdt, grid_box(o), mesh_box(o), w, wnew);
}

void oracle_addsource_grav(sac_oracle *o, double qdt, const double *w, double *wnew) {
  addsource_grav(o, qdt, mesh_box(o), w, wnew);
}

void oracle_centdiff4(sac_oracle *o, double qdt, const double *wCT, double *w) {
  centdiff4(o, qdt, grid_box(o), mesh_box(o), wCT, w);
}
