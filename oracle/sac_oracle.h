/*
 * sac_oracle.h -- CPU restatement of SAC's explicit perturbation-MHD update.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP library
 * in sac_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  The product path never calls it and has no CPU
 * fallback.
 *
 * The restatement follows the reference's Fortran array statements one by
 * one (index ranges, evaluation order, persistent scratch arrays), so that
 * quirks such as edge zeroing, the setnu mirror layer and stale scratch in
 * ghost cells behave as in the reference.  Each function cites the
 * reference file:line it restates.
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * of sac_oracle.c.
 */
#ifndef SAC_ORACLE_H
#define SAC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sac_oracle sac_oracle;

/* boundary type codes (typeB strings of src/vacgrid.t:257-464) */
enum {
  OR_BC_CONT = 1,
  OR_BC_FIXED = 2,
  OR_BC_PERIODIC = 3,
  OR_BC_SYMM = 4,
  OR_BC_SYMM0 = 5,
  OR_BC_ASYMM = 6,
  OR_BC_CONT1 = 7,
  OR_BC_FIXED1 = 8,
  OR_BC_GRAD1 = 9,
  OR_BC_CONTCD4 = 10,
  OR_BC_ZERO = 11
};

/* typeadvance codes (src/vac.t:406-480) */
enum {
  OR_ADV_ONESTEP = 1,
  OR_ADV_TWOSTEP = 2,
  OR_ADV_THREESTEP = 3,
  OR_ADV_FOURSTEP = 4,
  OR_ADV_STERCK = 5,
  OR_ADV_JAMESON = 6
};

/*
 * Create a problem.  nx[ndim] = mesh cells per axis (ghost width 2 is fixed,
 * src/vacdef.t:23).  eqpar[neqpar] is the reference's eqpar array with
 * gamma_=1, eta_=2 (src/vacpar.t.mhd:26) and the user parameters at the
 * 1-based slots grav0_ (gravity component d at grav0_+d) and nu_
 * (src/vacusrpar.f:3-5).  typeB[iw*2*ndim + iB] holds the boundary code of
 * variable iw (0-based) on boundary iB (0-based, lower/upper per axis,
 * src/vacgrid.t:53-69).  Returns NULL on bad input.
 */
sac_oracle *oracle_new(int ndim, int ndir, const int *nx, const double *eqpar,
                       int neqpar, int grav0_, int nu_, const int *typeB,
                       int fullgridini);
void oracle_free(sac_oracle *o);

/* grid extents */
int oracle_nw(const sac_oracle *o);
long oracle_ncell(const sac_oracle *o);

/* x[idim*ncell + cell], full grid (ghost entries used only if fullgridini) */
int oracle_set_x(sac_oracle *o, const double *x);
/* dx[idim*ncell + cell] after gridsetup1 */
int oracle_get_dx(const sac_oracle *o, double *dx);

/* state w[iw*ncell + cell] (Fortran column-major, index 1 fastest) */
int oracle_set_w(sac_oracle *o, const double *w);
int oracle_get_w(const sac_oracle *o, double *w);

/* iws list (1-based variable numbers); default all evolving variables */
int oracle_set_iws(sac_oracle *o, const int *iws, int niws);
/* Powell's div B source terms in addsource (divbfix, ndim > 1) */
void oracle_set_divbfix(sac_oracle *o, int on);
/* resistive sources (eta = eqpar(2) > 0): compactres selects addsource_res1,
 * else res2; dtdiffpar is the par value (getdt_visc later sets 0.25) */
void oracle_set_resistive(sac_oracle *o, int compactres, double dtdiffpar);
double oracle_getdt_res(sac_oracle *o, double dt);

/* getboundary(t,1,nw,1,ndim,w) on the state (src/vacgrid.t:195) */
void oracle_getboundary(sac_oracle *o);
/* getboundary on a caller array of the same layout */
void oracle_getboundary_on(sac_oracle *o, double *w);

/* getdt_courant (src/vac.t:641); returns dt, courantmax via out params */
double oracle_getdt_courant(sac_oracle *o, double courantpar,
                            double *courantmax);
/* getdt_special = getdt_visc (if nu on) + getdt_grav (src/vacusr.f:1413) */
double oracle_getdt_special(sac_oracle *o, double dt_in);
double oracle_maxviscoef(const sac_oracle *o);
void oracle_set_maxviscoef(sac_oracle *o, double v);

/* advance(iw_full,w) with typeadvance (src/vac.t:224,364) */
int oracle_advance(sac_oracle *o, double dt, int typeadvance);

/* advect1 with method 'cd4' on caller arrays: w <- w + qdt*R(wCT),
 * then getboundary(w) (src/vac.t:490) */
void oracle_advect1_on(sac_oracle *o, double qdt, const double *wCT,
                       double *w);

/* one main-loop iteration (src/vac.t:60-104): dt from courant/special,
 * then advance; returns dt used.  it/t bookkeeping in the oracle. */
double oracle_step(sac_oracle *o, double courantpar, int typeadvance);
void oracle_set_time(sac_oracle *o, int it, int itmin, double t);
double oracle_time(const sac_oracle *o);
int oracle_it(const sac_oracle *o);

/* building blocks on caller arrays (full grid, one field) */
void oracle_gradient4(sac_oracle *o, const double *q, int idim, double *gq);
void oracle_gradient1(sac_oracle *o, const double *q, int idim, int kind,
                      double *gq); /* kind 0 central, 1 L, 2 R (range mesh+1) */
void oracle_getflux(sac_oracle *o, const double *w, int iw, int idim,
                    double *f, int *transport); /* on mesh+2 along idim */
void oracle_getcmax(sac_oracle *o, const double *w, int idim, double *cmax);
/* setnu on the mesh; iw 1-based or 173 for the energy code */
void oracle_setnu(sac_oracle *o, const double *w, int iw, int idim,
                  double *nuR, double *nuL);
/* addsource_visc / addsource_grav: wnew += qdt*S(w) */
void oracle_addsource_visc(sac_oracle *o, double qdt, const double *w,
                           double *wnew);
void oracle_addsource_grav(sac_oracle *o, double qdt, const double *w,
                           double *wnew);
/* centdiff4 incl. addsource2 (src/vaccd.t:7), no getboundary */
void oracle_centdiff4(sac_oracle *o, double qdt, const double *wCT,
                      double *w);

/* fill every persistent scratch array with a value (for tests showing that
 * mesh results do not depend on stale scratch) */
void oracle_poison_scratch(sac_oracle *o, double value);

#ifdef __cplusplus
}
#endif
#endif
