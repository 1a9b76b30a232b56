/*
 * mercury_oracle.h -- CPU restatement of MERCURY6's massless-body hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path in
 * mercuda_amd/ is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function restates one Fortran routine of mercury6_2.for (cited per
 * function in mercury_oracle.c) in the same floating-point operation order, so
 * that a build with -ffp-contract=off reproduces the reference bit for bit
 * except where the reference calls libm (cos, pow, fmod).
 *
 * Array conventions follow the Fortran: x[3*j+k] is component k of body j,
 * body 0 is the central body, bodies 1..nbig-1 are Big, nbig..nbod-1 Small.
 * Masses are G*M (AU^3/day^2).  Indices returned in event lists are 1-based,
 * as in the Fortran.
 */
#ifndef MERCURY_ORACLE_H
#define MERCURY_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kepler drift chain: drift_one -> drift_dan -> drift_kepmd / drift_kepu. */
int orc_drift_one(double mu, double *xv, double dt);          /* xv[6] in/out, returns iflg */
int orc_drift_dan(double mu, double *xv, double dt);
void orc_drift_kepu_stumpff(double x, double *c);              /* c[4] = c0..c3 */
void orc_mco_sine(double x, double *sx, double *cx);
/* Which branch drift_dan takes for this state: 0 kepmd, 1 kepu (Newton), 2
 * kepu (Laguerre), 3 failed (iflg != 0).  Used only to label test vectors. */
int orc_drift_branch(double mu, const double *xv, double dt);

/* Coordinate transforms (Big bodies only, as mdt_mvs calls them). */
void orc_mco_h2j(int nbig, const double *m, const double *xh, const double *vh,
                 double *xj, double *vj);
void orc_mco_j2h(int nbig, const double *m, const double *xj, const double *vj,
                 double *xh, double *vh);

/* Forces. */
void orc_mfo_obl(const double *jcen, int nbod, const double *m, const double *x,
                 double *a, double *acen);
void orc_mfo_mvs(const double *jcen, int nbod, int nbig, const double *m,
                 const double *x, const double *xj, double *a);

/* Encounter screening. */
void orc_mce_min(double d0, double d1, double d0t, double d1t, double h,
                 double *d2min, double *tmin);
void orc_mce_box(int nbod, double h, const double *x0, const double *v0,
                 const double *x1, const double *v1, double *xmin, double *xmax,
                 double *ymin, double *ymax);

typedef struct {
    int cap;                       /* capacity of every list below */
    int nclo;                      /* close-encounter minima */
    int *iclo, *jclo;
    double *dclo, *tclo, *ixvclo, *jxvclo;   /* ixvclo[6*k+c] */
    int nhit;                      /* near misses / collisions */
    int *ihit, *jhit, *chit;
    double *dhit, *thit;
    double thit1;
    int nowflag;
    int overflow;                  /* set when a list would exceed cap */
} orc_events;

void orc_mce_stat(double time, double h, double rcen, int nbod, int nbig,
                  const double *m, const double *x0, const double *v0,
                  const double *x1, const double *v1, const double *rce,
                  const double *rphys, orc_events *ev);
/* Central-body impacts over bodies i0..nbod (1-based i0, as the Fortran). */
int orc_mce_cent(double time, double h, double rcen, int i0, int nbod,
                 const double *m, const double *x0, const double *v0,
                 const double *x1, const double *v1, int cap, int *jhit,
                 double *thit, double *dhit);

/* One MVS step of the whole system (mdt_mvs).  a[3*nbod] carries the
 * accelerations between calls; *dtflag != 2 recomputes them first. */
void orc_mdt_mvs(double time, double h0, double rcen, const double *jcen,
                 int nbod, int nbig, const double *m, double *x, double *v,
                 double *a, int *dtflag, const double *rce, const double *rphys,
                 orc_events *ev);

/* Symplectic corrector mco_mvs2h: symplectic (x,v) -> physical (xh,vh). */
void orc_mco_mvs2h(double h, const double *jcen, int nbod, int nbig,
                   const double *m, const double *x, const double *v,
                   double *xh, double *vh);

/* The non-gravitational model (ngflag, ngf[5*nbod]) of the calls that follow;
 * orc_set_ngf(0, NULL, 0) clears it.  ngf stays the caller's. */
void orc_set_ngf(int ngflag, const double *ngf, int nbod);
/* mfo_ngf (cometary A1..A3, Yarkovsky) and mfo_pr (radiation pressure, PR
 * drag) for every body; orc_mfo_ngf returns 1 for an undefined direction. */
int orc_mfo_ngf(int nbod, const double *x, const double *v, double *a, const double *ngf);
void orc_mfo_pr(int nbod, const double *m, const double *x, const double *v, double *a, const double *ngf);
/* Heliocentric accelerations of mfo_all without user forces: mfo_grav (pn = 1
 * adds the central 1PN term) + mfo_obl + the orc_set_ngf model. */
void orc_mfo_all(const double *jcen, int nbod, int nbig, const double *m,
                 const double *x, const double *v, int pn, double *a);

/* RA15 state kept between calls (mdt_ra15's saved b, e). */
typedef struct {
    int nbod;
    int dtflag;           /* 0 first call, 1 reset, 2 normal (as mdt_ra15) */
    double *b, *e;        /* [3*nbod][7] */
} orc_ra15;

/* One RA15 sequence (mdt_ra15 mercury6_2.for:4133-4500) with mfo_all forces.
 * *t in: trial sequence size, out: proposal for the next call; *tdid: size
 * done.  Returns the number of force evaluations; *redo counts the
 * shrink-and-redo passes of a reset call. */
long orc_mdt_ra15(double time, double *t, double *tdid, double tol, const double *jcen,
                  int nbod, int nbig, const double *m, double *x1, double *v1, int pn,
                  orc_ra15 *st, int *redo);

/* nsteps RA15 sequences starting from trial size h0; hdid[] receives each
 * size done.  Returns force evaluations. */
long orc_ra15_run(double time0, double h0, double tol, const double *jcen, int nbod,
                  int nbig, const double *m, double *x, double *v, int pn, int nsteps,
                  double *hdid, double *hnext);

/* ---- HYBRID (algorithm 10) ---- */
/* Interaction accelerations mfo_hy: changeover-weighted direct terms + oblateness. */
void orc_mfo_hy(const double *jcen, int nbod, int nbig, const double *m, const double *x,
                const double *rcrit, double *a);
/* mce_snif: encounter pairs (1-based, loop order) of one step; -1 on overflow. */
int orc_mce_snif(double h, int nbod, int nbig, const double *x0, const double *v0, const double *x1,
                 const double *v1, const double *rcrit, int *ce, int cap, int *ice, int *jce);
/* mfo_hkce: the encounter set's Keplerian-part accelerations (listed pair
 * terms weighted 1 - K, solar terms); pairs 1-based. */
void orc_mfo_hkce(int nbod, const double *m, const double *x, const double *rcrit, double *a, int nce,
                  const int *ice, const int *jce);
/* mdt_bs2 with mfo_hkce forces on a (compact) set; pairs 1-based. */
int orc_mdt_bs2(double time, double *h0, double *hdid, double tol, int nbod, const double *m, double *x0,
                double *v0, const double *rcrit, int nce, const int *ice, const int *jce, long *nforce);
/* algorithm 2 (mdt_bs1) or 3 (mdt_bs2) with mfo_all: one step; a run of steps */
int orc_mdt_bs(int algor, double time, double *h0, double *hdid, double tol, const double *jcen, int nbod,
               int nbig, const double *m, double *x0, double *v0, int pn, long *nforce);
long orc_bs_run(int algor, double time0, double h0, double tol, const double *jcen, int nbod, int nbig,
                const double *m, double *x, double *v, int pn, int nsteps, double *hdid, double *hnext,
                long *rejected);
typedef struct {
    double *a;            /* interaction accelerations carried between calls */
    int dtflag;           /* 0 first call, 1 reset, 2 normal */
    double hrec;          /* BS2 step recommendation kept between encounters */
    long nforce;
    int cap;              /* capacity of ice/jce */
    int *ice, *jce;
} orc_hybrid;
/* One HYBRID step (mdt_hy, collisions not allowed); events of the encounter
 * substeps in ev (close-encounter list, global 1-based indices).  Returns
 * the number of encounter pairs, -1 on failure. */
int orc_mdt_hy(double time, double h0, double tol, const double *jcen, double rcen, int nbod, int nbig,
               const double *m, double *x, double *v, const double *rcrit, const double *rce,
               const double *rphys, orc_hybrid *st, orc_events *ev);
void orc_mco_dh2h(int nbod, const double *m, const double *x, const double *v, double *xh, double *vh);
/* nsteps HYBRID steps as the driver takes them, with mce_cent on the
 * physical states after each (rows as orc_mvs_run). */
int orc_hy_run(double time0, double h0, double tol, double rcen, const double *jcen, int nbod, int nbig,
               const double *m, double *x, double *v, const double *rcrit, const double *rce,
               const double *rphys, int nsteps, int cap, int *evi, double *evd, long *nforce);

/* Convenience for tests and the CPU baseline: nsteps MVS steps of a whole
 * system with the driver's central-impact check after each, event lists
 * flattened into caller arrays.  Returns the number of events written
 * (or -1 if cap was exceeded).  Each event row: step, kind, i, j, and
 * (distance, time) in evd[2*k]. */
int orc_mvs_run(double time0, double h0, double rcen, const double *jcen,
                int nbod, int nbig, const double *m, double *x, double *v,
                const double *rce, const double *rphys, int nsteps,
                int screen, int cap, int *evi, double *evd);

#ifdef __cplusplus
}
#endif
#endif
