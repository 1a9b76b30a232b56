/*
 * mercury_hip.h -- C ABI of libmercury_hip, the MI355X (gfx950) backend of the
 * MERCURY6 massless-body hot path.
 *
 * The first eighteen entry points replace, one for one and with the same
 * signatures and semantics, the CUDA backend's ABI declared in the reference's
 * mercury_cuda.h (cited per function below); only the prefix changes from
 * mercury_cuda_ to mercury_hip_, because the Fortran binding over them
 * (mercury_gpu.f90) is rewritten for this library.  Arrays at the boundary keep
 * MERCURY's Fortran layout: x(3,n), v(3,n) with body 1 the central body,
 * masses and jcen pre-multiplied by K2.  Every int return is 0 on success and
 * nonzero on failure with a "MERCURY HIP: ..." line on stderr; no exception
 * crosses the ABI.  One global context per process; not re-entrant.
 *
 * The entry points after them have no counterpart in the reference: fused
 * multi-step advance, batched ensembles and the tangent-linear/adjoint step.
 */
#ifndef MERCURY_HIP_H
#define MERCURY_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as the reference's struct MercuryEvent (mercury_cuda.h:7-11):
 * 1-based body indices i, j; kind bits 1 close-encounter minimum, 2 near miss,
 * 4 collision, 8 central-body impact; now = separation at step end within the
 * collision distance; xi, xj = states interpolated to the minimum. */
typedef struct MercuryEvent {
    int i, j, kind, now;
    double distance, time;
    double xi[6], xj[6];
} MercuryEvent;

/* ---- the reference ABI (mercury_cuda.h:12-34) --------------------------- */

/* mercury_cuda.h:13  1 if a HIP device is present, else 0. */
int mercury_hip_available(void);
/* mercury_cuda.h:14  select the algorithm: 1 MVS, 2 BS, 3 BS2, 4 RADAU,
 * 9 TEST (MVS, no corrector), 10 HYBRID; any other returns 1.  Every
 * algorithm takes up to 63 Big bodies besides the central body. */
int mercury_hip_configure(int algorithm);
/* mercury_cuda.h:15  flag != 2 discards the MVS acceleration / RADAU history. */
void mercury_hip_reset(int flag);
/* mercury_cuda.h:16  new close-encounter limits rce[n]; history is kept. */
int mercury_hip_limits(const double *rce);
/* mercury_cuda.h:17  copy out the state; physical = 1 applies the MVS
 * symplectic corrector (mco_mvs2h) on scratch, or gives HYBRID's physical
 * velocities. */
int mercury_hip_export(double h, int physical, double *x, double *v);
/* mercury_cuda.h:18-21  load a system: n bodies, nbig Big (central included),
 * pn = 1PN flag, ngflag = non-gravitational flag, m[n], x[3n], v[3n],
 * ngf[5n], jcen[3], rce[n], rphys[n]. */
int mercury_hip_upload(int n, int nbig, int pn, int ngflag,
                       const double *m, const double *x, const double *v,
                       const double *ngf, const double *jcen, const double *rce,
                       const double *rphys);
/* mercury_cuda.h:22-23  one step from time; MVS: *hdid = *h; RADAU: *h is
 * replaced by the next proposal.  forces/rejected receive counts. */
int mercury_hip_step(double time, double *h, double *hdid, double tol,
                     int64_t *forces, int64_t *rejected);
/* mercury_cuda.h:24  state after the last step, or before it (previous = 1). */
int mercury_hip_download(double *x, double *v, int previous);
/* mercury_cuda.h:25-26  events of the last step, library-owned array valid
 * until the next call, sorted by (min, max) body index. */
int mercury_hip_events(double time, double h, double rcen,
                       const MercuryEvent **events, int *count);
/* mercury_cuda.h:27  heliocentric accelerations a[3n] of the uploaded state. */
int mercury_hip_force(double *a);
/* mercury_cuda.h:28  HYBRID: kick, solar drift, Kepler drift, changeover scan. */
int mercury_hip_hybrid_begin(double h, const double *crit, int flag, int cap, int *ce,
                             int *nce, int *pi, int *pj, double *x, double *v,
                             int64_t *forces);
/* mercury_cuda.h:29  HYBRID: scatter members, solar drift, force, kick. */
int mercury_hip_hybrid_finish(double h, const double *m, const double *x,
                              const double *v, int64_t *forces);
/* mercury_cuda.h:30  HYBRID compact encounter context: enter. */
int mercury_hip_encounter_enter(int n, int nbig, const double *m, const double *x,
                                const double *v, const double *crit, const double *rce,
                                const double *rphys, int np, const int *pi, const int *pj);
/* mercury_cuda.h:31  HYBRID compact encounter context: replace m, x, v. */
int mercury_hip_encounter_update(const double *m, const double *x, const double *v);
/* mercury_cuda.h:32  HYBRID compact encounter context: leave. */
void mercury_hip_encounter_exit(void);
/* mercury_cuda.h:33  release every device allocation. */
void mercury_hip_free(void);

/* ---- extensions (no counterpart in the reference) ----------------------- */

/* Event of a multi-step or batched launch: step within the call, system. */
typedef struct MercuryHipEvent {
    MercuryEvent e;
    int step, system;
} MercuryHipEvent;

/* nsteps MVS steps of the uploaded system in fused launches (steps_per_launch
 * steps each), with in-kernel encounter/central screening when screen != 0.
 * Same per-step arithmetic as mercury_hip_step. */
int mercury_hip_advance(double time, double h, int nsteps, double rcen, int screen,
                        int64_t *forces);
/* Batched ensemble: nsys independent systems in one context.  Arrays are the
 * systems' MERCURY-layout arrays concatenated (system s has n[s] bodies, its
 * first the central body, nbig[s] Big bodies including it); h[s] and rcen[s]
 * are per system.  Every launch of the advance covers all systems (planet
 * tables per system, workgroups never spanning two systems).  Replaces the
 * single-system state; free/configure/upload as for one system. */
int mercury_hip_ensemble_upload(int nsys, const int *n, const int *nbig, const double *m,
                                const double *x, const double *v, const double *rce,
                                const double *rphys, const double *h, const double *rcen);
/* nsteps MVS steps of every system of the ensemble from `time`. */
int mercury_hip_ensemble_advance(double time, int nsteps, int screen, int64_t *forces);
/* States of all systems, concatenated as uploaded. */
int mercury_hip_ensemble_download(double *x, double *v);
/* Events of the last mercury_hip_advance / ensemble advance, sorted by
 * (system, step, min index, max index). */
int mercury_hip_advance_events(const MercuryHipEvent **events, int *count);
/* Launch shape: steps fused per particle launch (>= 1), particle workgroup
 * size (64..256, multiple of 64), particles per lane of the fast particle
 * kernel (1..8; each lane takes that many particles of its tile one after the
 * other).  0 keeps the current value. */
int mercury_hip_set_tuning(int steps_per_launch, int block_size, int particles_per_lane);
/* Kernel timing with HIP events on the launch streams: enable = 1 clears and
 * starts recording every kernel kind, 2 the fast particle kernel only (no
 * events on the planet and slow-set streams), 0 stops; the read returns the
 * summed milliseconds and launch counts of the fast particle kernel, the
 * planet kernel and the slow-set kernel.  RADAU's particle kernel counts as
 * the particle kernel. */
int mercury_hip_timing(int enable);
int mercury_hip_timing_read(double *particle_ms, int *particle_launches,
                            double *planet_ms, int *planet_launches,
                            double *general_ms, int *general_launches);
/* Particles parked into the general kernel by the last advance. */
int64_t mercury_hip_parked(void);
/* Particle-steps completed by the fast and by the general particle kernel
 * since timing was last enabled. */
int mercury_hip_work_read(int64_t *fast_steps, int64_t *general_steps);

/* The driver's encounter loop mdt_hkce (mercury6_2.for:3883-3905) over one step
 * h0 in the active encounter context, with mdt_bs2's step control on the device
 * (no counterpart in the reference, whose driver calls mercury_hip_step once per
 * substep): for callers that do not list encounters after every substep.  hrec:
 * in, the BS step recommendation kept from the last encounter; out, the new
 * one.  Returns the substeps taken, force evaluations and rejected attempts;
 * the states are those of the per-substep calls, bit for bit. */
int mercury_hip_encounter_advance(double h0, double *hrec, double tol, int64_t *substeps,
                                  int64_t *forces, int64_t *rejected);

/* Device self-check of the arithmetic (no counterpart in the reference): the
 * fast MVS kernel divides and takes square roots with shorter sequences that
 * are exact for operands inside the normal range.  This draws n random
 * operand sets from `seed` over those ranges and writes, for the divide, the
 * reciprocal, the square root and the divide by a constant, how many results
 * differ in any bit from the compiler's IEEE expansions (mismatches[4]). */
int mercury_hip_check_arith(int64_t n, uint64_t seed, int64_t *mismatches);

/* hipGraph mode for mercury_hip_advance / ensemble advance (no counterpart in
 * the reference): enable = 1 captures a call's whole launch sequence (all
 * streams) into a graph and replays it for later calls with the same step
 * count, flags, tuning and buffers; results are bitwise those of the eager
 * launches.  Off by default; ignored while kernel timing is on. */
int mercury_hip_set_graph(int enable);
/* Graphs captured and replays made since the library was loaded. */
int mercury_hip_graph_stats(int *captures, int *replays);
/* Autotuner (no counterpart in the reference): times probe_steps MVS steps of
 * the uploaded system or ensemble for every steps-per-launch in {8,16,32,64},
 * workgroup size in {64,128,256} and particles per lane in {1,2,4} (h: the
 * step of a single system; ensemble members use their own), keeps the fastest
 * (mercury_hip_set_tuning) and leaves the state as it was.  Returns the choice
 * and its particle-steps/s. */
int mercury_hip_autotune(double h, int probe_steps, int *steps_per_launch, int *block_size,
                         int *particles_per_lane, double *particle_steps_per_s);

/* Allocation-free force (no counterpart in the reference): for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The accelerations mercury_hip_force
 * gives (mfo_all: gravity, 1PN when pn != 0, oblateness from jcen), for a
 * system passed in DEVICE memory in this library's layout: m[n]; x, v and the
 * output a component-major, x[c*ld + j] for c = 0..2 (v is read only when
 * pn != 0); jcen[3] is host memory.  Body 0 is the central body (its row of a
 * is zero).  Runs on the null stream and returns when a is written.  It
 * allocates nothing: its scratch is a fixed 4 KB device array per device,
 * which holds at most 64 Big bodies (more is an error), and calls that
 * overlap on one device race on it. */
int mercury_hip_force_noalloc(int n, int nbig, int pn, const double *m, const double *x, const double *v,
                              const double *jcen, double *a, long long ld);

/* ---- Differentiation of the MVS map (no counterpart in the reference) ----
 * For orbit fitting: gradients of a function of the final states with
 * respect to the initial states of every body and the Big-body masses.
 * Single uploaded MVS system, jcen = 0, at most 12 Big bodies besides the
 * central body.  The uploaded state is left unchanged.
 *
 * Forward pass of nsteps MVS steps of size h from the uploaded state, keeping
 * a tape: the planet tables of every step and the particles' states every
 * `segment` steps (0: ceil(sqrt(nsteps))).  Writes the final states x[3n],
 * v[3n] (the same numbers mercury_hip_advance produces). */
int mercury_hip_adjoint_forward(int nsteps, double h, int segment, double *x, double *v);
/* Reverse pass over the tape: given cotangents gx[3n], gv[3n] of the final
 * states (particles and Big bodies; the central body's heliocentric state is
 * identically zero, so its rows contribute nothing), the gradient with
 * respect to the initial states (gx0[3n], gv0[3n]) and the masses (gm[n]; Big
 * bodies only, zero elsewhere). */
int mercury_hip_adjoint_backward(const double *gx, const double *gv, double *gx0, double *gv0,
                                 double *gm);
/* Tangent-linear pass over the tape: the derivative of the final states
 * along the direction (dx0[3n], dv0[3n], dm[n]) of initial states and
 * Big-body masses, into dx[3n], dv[3n]. */
int mercury_hip_tangent(const double *dx0, const double *dv0, const double *dm, double *dx, double *dv);

#ifdef __cplusplus
}
#endif
#endif
