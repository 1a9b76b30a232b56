"""Throughput of the MVS massless-particle step on one MI355X.

Workload: 8 planets + 1,000,000 massless test particles (seeded synthetic
system: planet masses log-U[1e-7,1e-3] Msun, a log-U[0.4,30] AU, >= 10 mutual
Hill radii apart; particles a U[0.5,40] AU, e U[0,0.3], i U[0,20 deg]),
FP64, step 4 d, encounter and central-impact screening on every step.  A
"step" is one MVS step of the whole system (planets and all particles):
kick, Kepler drift, close-encounter screen, planet-to-particle force, kick,
central-body check.  Timed region: `steps` steps through
mercury_hip_advance, state resident in HBM, host wall clock around the call
(which synchronises), after `warmup` untimed steps.

Prints one JSON line.  `roofline` refers to the dominant kernel (the fast
particle kernel) timed with HIP events on its launch stream; `cpu_baseline` is
the reference's own CPU path (mdt_mvs + mce_cent from oracle/_ref, compiled
from the reference Fortran) on a bounded sample, one process per core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Algorithmic FP64 operations per fast-path particle-step (kick, kepmd drift,
# particle box + 8 planet box tests, 8-planet force, kick, central gate),
# frozen from `python -m oracle.opcount` over this workload's population
# (each IEEE +,-,*,/,sqrt = 1; no FMA in this build).
FLOP_PER_PSTEP = 351.6
# Algorithmic HBM bytes per particle per launch of the fast kernel: read
# x, v, a (72 B) and rce (8 B), write x, v, a (72 B).
BYTES_PER_PARTICLE_LAUNCH = 152.0
# HBM bytes per particle per launch of the fast kernel as rocprofv3 measured
# them on this workload (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, the factors
# a 256 MiB device copy reads on the same counters; 1M particles, the 63
# 32-step launches of the default timed run): profiles/mvs_particles_fast_1M_hbm_traffic.md.
# Beyond the state it includes the screening requests the kernel appends.
TRAFFIC_PER_PARTICLE_LAUNCH = 177.6
# Algorithmic FP64 operations per particle per steady RA15 sequence (2
# iterations, 15 force evaluations; tol 1e-12, central 1PN, 8 planets), from
# the counted restatement of mdt_ra15 checked bit for bit against the oracle
# (`python -m oracle.opcount`; oracle.opcount.ra15_measure).
FLOP_PER_RA15_PSEQ = 5550.0
# Algorithmic HBM bytes per particle per sequence of the RA15 particle kernel:
# read x, v, b, e, a1 (51 doubles), write them back (51 doubles).
BYTES_PER_RA15_PSEQ = 816.0
# HBM bytes per particle per sequence of the RA15 particle kernel as rocprofv3
# measured them (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, 1M particles, the
# steady sequences): profiles/ra15_1M_kernel_trace.md.
TRAFFIC_PER_RA15_PSEQ = 882.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(sysd, nproc=16, shard=31250, nsteps=1536):
    """The reference CPU path (flang build of mdt_mvs + mce_cent) on `nproc`
    disjoint particle shards in parallel processes.  Massless particles do not
    interact, so the shards are exact.  Returns particle-steps/s."""
    import multiprocessing as mp
    import numpy as np
    from oracle.bindings import REF_PATH
    if not REF_PATH.exists():
        return None
    nb = sysd["nbig"]
    rng = np.random.default_rng(1)
    pick = rng.choice(np.arange(nb, len(sysd["m"])), nproc * shard, replace=False)
    jobs = []
    for p in range(nproc):
        idx = np.concatenate([np.arange(nb), pick[p * shard:(p + 1) * shard]])
        sub = {k: sysd[k][idx] for k in ("m", "x", "v", "rce", "rphys")}
        sub.update(nbig=nb, h=sysd["h"], rcen=sysd["rcen"], jcen=np.zeros(3))
        jobs.append((sub, nsteps))
    ctx = mp.get_context("spawn")
    with ctx.Pool(nproc) as pool:
        times = pool.map(_cpu_worker, jobs)
    return nproc * shard * nsteps / max(times), max(times)


def _cpu_worker(job):
    import threading
    sub, nsteps = job
    out = {}

    def run():
        from oracle.bindings import Reference
        R = Reference(nmax=len(sub["m"]) + 16, cmax=len(sub["m"]) + 16)
        R.mdt_mvs_run(sub, 2)           # warm (allocations, page-in)
        t = time.perf_counter()
        R.mdt_mvs_run(sub, nsteps)
        out["t"] = time.perf_counter() - t

    threading.stack_size(256 << 20)
    th = threading.Thread(target=run)
    th.start()
    th.join()
    return out["t"]


def _in_big_thread(fn):
    """Run fn in a thread with a large stack (the Fortran keeps big locals)."""
    import threading
    out = {}

    def run():
        out["r"] = fn()

    threading.stack_size(256 << 20)
    th = threading.Thread(target=run)
    th.start()
    th.join()
    return out["r"]


def _cpu_ra15_worker(job):
    sub, nseq = job

    def run():
        from oracle.bindings import Reference
        R = Reference(nmax=len(sub["m"]) + 16, cmax=len(sub["m"]) + 16)
        R.ra15_run(sub, 1, pn=1)                       # warm
        t = time.perf_counter()
        R.ra15_run(sub, nseq, pn=1)
        return time.perf_counter() - t
    return _in_big_thread(run)


def cpu_radau(sysd, nproc=16, shard=62500, nseq=8):
    """The reference mdt_ra15 with mfo_all (tol 1e-12, 1PN) on `nproc` disjoint
    particle shards of the 1M system in parallel processes.  Each shard takes
    its own step sizes (the RA15 law uses max |b7| over the bodies it
    integrates), close to the whole system's.  Returns particle-steps/s."""
    import multiprocessing as mp
    import numpy as np
    from oracle.bindings import REF_PATH
    if not REF_PATH.exists():
        return None
    nb = sysd["nbig"]
    jobs = []
    for p in range(nproc):
        idx = np.concatenate([np.arange(nb), nb + p * shard + np.arange(shard)])
        sub = {k: sysd[k][idx] for k in ("m", "x", "v", "rce", "rphys")}
        sub.update(nbig=nb, h=sysd["h"], rcen=sysd["rcen"], jcen=np.zeros(3))
        jobs.append((sub, nseq))
    with mp.get_context("spawn").Pool(nproc) as pool:
        times = pool.map(_cpu_ra15_worker, jobs)
    return nproc * shard * nseq / max(times), max(times)


def cpu_hybrid(nparticles=20_000, nsteps=1):
    """The reference mdt_hy (with its mdt_hkce BS2 encounter loop) on the same
    20k-particle encounter system as the GPU's hybrid_rate(20000), one process:
    the reference integrates one system serially, and splitting its particles
    would change the encounter set that shares a BS2 step."""
    from oracle.bindings import REF_PATH, Reference
    from mercuda_amd import systems
    if not REF_PATH.exists():
        return None
    s = systems.config_encounters(seed=3, nparticles=nparticles)

    def run():
        R = Reference(nmax=len(s["m"]) + 16, cmax=len(s["m"]) * 8 + 16)
        t = time.perf_counter()
        R.mdt_hy_run(s, nsteps)
        return time.perf_counter() - t
    sec = _in_big_thread(run)
    return nparticles * nsteps / sec, sec


def ensemble_ratio(hip, steps, spl, block, replays=7):
    """Batched ensemble (512 systems, 2..12 planets, log-U[100, 5e4] particles,
    dt 4 d) against one 8-planet system with the same total particle count,
    both through mercury_hip_advance with screening on.  Each side is
    autotuned (steps per launch, workgroup size, particles per lane) on its
    own state and runs as a replayed hipGraph; each side's rate is the median
    of `replays` timed replays of `steps` steps (min and max reported)."""
    from mercuda_amd import systems
    log("ensemble: generating 512 systems")
    ens = systems.ensemble(seed=4)
    ntot = sum(len(q["m"]) - q["nbig"] for q in ens)

    def timed(advance):
        advance(0.0, 16)                          # first call (acceleration reset)
        advance(16 * 4.0, steps)                  # captures the steady shape
        hip.graph_stats()
        secs = []
        for r in range(replays):
            tic = time.perf_counter()
            advance((16 + (r + 1) * steps) * 4.0, steps)        # replay
            secs.append(time.perf_counter() - tic)
        return sorted(secs)

    hip.configure(1)
    hip.ensemble_upload(ens)
    hip.set_tuning(steps_per_launch=spl, block_size=block)
    tune_e = hip.autotune(probe_steps=64)
    hip.set_graph(True)
    t_ens = timed(lambda t, n: hip.ensemble_advance(t, n, screen=True))
    hip.set_graph(False)
    log(f"ensemble: single system with {ntot} particles")
    one = systems.config_large(seed=2, nparticles=ntot)
    hip.configure(1)
    hip.upload_system(one)
    hip.set_tuning(steps_per_launch=spl, block_size=block)
    tune_1 = hip.autotune(h=one["h"], probe_steps=64)
    hip.set_graph(True)
    t_one = timed(lambda t, n: hip.advance(t, one["h"], n, one["rcen"], screen=True))
    hip.set_graph(False)
    med = replays // 2
    r_ens = ntot * steps / t_ens[med]
    r_one = ntot * steps / t_one[med]
    # spread of the ratio: slowest ensemble replay over fastest single one, and the reverse
    lo = t_one[0] / t_ens[-1]
    hi = t_one[-1] / t_ens[0]
    return {"systems": len(ens), "particles": ntot, "steps_per_replay": steps, "replays": replays,
            "ensemble_particle_steps_per_s": round(r_ens, 1),
            "single_system_particle_steps_per_s": round(r_one, 1),
            "ratio": round(r_ens / r_one, 3),
            "ratio_range": [round(lo, 3), round(hi, 3)],
            "ensemble_replay_ms": [round(1e3 * t, 3) for t in t_ens],
            "single_replay_ms": [round(1e3 * t, 3) for t in t_one],
            "ensemble_tuning": {k: tune_e[k] for k in ("steps_per_launch", "block", "particles_per_lane")},
            "single_tuning": {k: tune_1[k] for k in ("steps_per_launch", "block", "particles_per_lane")},
            "graph": hip.graph_stats()}


def adjoint_cost(hip, nsteps=10000, nparticles=4096):
    """Orbit-fitting workload: 8 planets + 4,096 massless particles, 10^4 MVS
    steps of 4 d, loss = sum of squared final position offsets to targets
    (initial positions + N(0, 1e-3 AU)).  Forward = the plain MVS integration
    (mercury_hip_advance, no screening); adjoint = the taped forward pass plus
    the reverse pass giving gradients for every initial state and the 8
    masses."""
    import numpy as np
    from mercuda_amd import systems
    s = systems.make_system(7, nparticles, nplanets=8, a_range=(0.5, 40.0), emax=0.3, imax_deg=20.0, h=4.0)
    s["jcen"] = np.zeros(3)
    nb = s["nbig"]
    target = s["x"][nb:] + np.random.default_rng(11).normal(0, 1e-3, (nparticles, 3))
    h = s["h"]
    hip.configure(1)
    hip.upload_system(s)
    hip.advance(0.0, h, 256, s["rcen"], screen=False)          # warm-up
    hip.upload_system(s)
    tic = time.perf_counter()
    hip.advance(0.0, h, nsteps, s["rcen"], screen=False)
    t_fwd = time.perf_counter() - tic
    hip.upload_system(s)
    xT, vT = hip.adjoint_forward(256, h)                      # warm-up
    hip.adjoint_backward(np.zeros_like(xT), np.zeros_like(vT))
    tic = time.perf_counter()
    xT, vT = hip.adjoint_forward(nsteps, h)
    gx = np.zeros_like(xT)
    gx[nb:] = 2.0 * (xT[nb:] - target)
    gx0, gv0, gm = hip.adjoint_backward(gx, np.zeros_like(vT))
    t_adj = time.perf_counter() - tic
    return {"workload": f"MVS, 8 planets + {nparticles} particles, {nsteps} steps of 4 d; gradients of "
                        "sum |x_T - target|^2 w.r.t. all initial states and the 8 masses",
            "forward_s": round(t_fwd, 4), "adjoint_s": round(t_adj, 4),
            "adjoint_over_forward": round(t_adj / t_fwd, 3),
            "checkpoint_interval": int(np.ceil(np.sqrt(nsteps)))}


def radau_rate(hip, nparticles=1_000_000, nseq=12):
    """RA15 (tol 1e-12, central 1PN) on 8 planets + 1M particles of the MVS
    workload's generator, first trial step 4 d: accepted sequences x
    particles per second through mercury_hip_step (one host round trip per
    sequence for the step-size law), after 3 warm-up sequences."""
    from mercuda_amd import systems
    s = systems.config_large(seed=2, nparticles=nparticles)
    hip.configure(4)
    hip.upload_system(s, pn=1)
    h, t = s["h"], 0.0
    for _ in range(3):
        h, hd, _, _ = hip.step(t, h, 1e-12)
        t += hd
    forces = 0
    days = 0.0
    hip.timing(True, fast_only=True)      # HIP events around the particle kernel's launches
    tic = time.perf_counter()
    for _ in range(nseq):
        h, hd, nf, _ = hip.step(t, h, 1e-12)
        t += hd
        days += hd
        forces += nf
    sec = time.perf_counter() - tic
    tm = hip.timing_read()
    hip.timing(False)
    launches = max(tm["particle_launches"], 1)
    avg_s = tm["particle_ms"] / launches / 1e3
    # one particle-kernel launch per steady sequence (a redo adds one)
    flop = FLOP_PER_RA15_PSEQ * nparticles * nseq / launches
    roofline = {"bound": "fp64_valu", "kernel": "ra_particles", "achieved": round(flop / avg_s / 1e12, 3),
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(flop / avg_s / 1e12 / PEAK_FP64_TFLOPS, 4),
                "traffic": None if TRAFFIC_PER_RA15_PSEQ is None else round(TRAFFIC_PER_RA15_PSEQ * nparticles),
                "avg_launch_ms": round(avg_s * 1e3, 4), "launches": tm["particle_launches"],
                "flop_per_particle_sequence": FLOP_PER_RA15_PSEQ,
                "hbm_bytes_algorithmic_per_launch": round(BYTES_PER_RA15_PSEQ * nparticles)}
    return {"workload": f"RADAU tol 1e-12 with 1PN, 8 planets + {nparticles} particles",
            "sequences": nseq, "particle_steps_per_s": round(nparticles * nseq / sec, 1),
            "particle_force_evaluations_per_s": round(nparticles * forces / sec, 1),
            "ms_per_sequence": round(sec * 1e3 / nseq, 3), "mean_step_days": round(days / nseq, 4),
            "roofline": roofline}


PUBLISHED_EXTRA_FORCES = {"cuda_rtx4070_s": 6.04, "mercuda_cpu_1thread_s": 33.4, "mercury6_matched_s": 90.9,
                          "source": "reference docs/benchmarks.md:36-44 (whole program: startup, integration, "
                                    "final output)"}


def radau_extra_forces(hip, nparticles=100_000, days=365.0, tol=1e-11):
    """The reference's published RADAU case with extra forces: solar 1PN,
    radiation pressure and PR drag (beta 1e-4) and Yarkovsky (1e-12 AU/day^2)
    on 8 planets + 100k particles for 365 days, tol 1e-11, first step 1 d
    (systems.config_extra_forces).  Wall time of upload, the mercury_hip_step
    loop to day 365 (the last size clipped to it) and the final download; the
    reference's figure also counts its program start and file output."""
    from mercuda_amd import systems
    s = systems.config_extra_forces(nparticles=nparticles)
    hip.configure(4)
    hip.upload_system(dict(s, x=s["x"][:s["nbig"] + 1], v=s["v"][:s["nbig"] + 1], m=s["m"][:s["nbig"] + 1],
                           rce=s["rce"][:s["nbig"] + 1], rphys=s["rphys"][:s["nbig"] + 1],
                           ngf=s["ngf"][:s["nbig"] + 1]), pn=1)
    hip.step(0.0, 1.0, tol)                      # warm: kernels loaded, arrays sized
    tic = time.perf_counter()
    hip.upload_system(s, pn=1)
    h, t, nseq, forces = s["h"], 0.0, 0, 0
    while t < days:
        hn, hd, nf, _ = hip.step(t, min(h, days - t), tol)
        t += hd
        nseq += 1
        forces += nf
        h = hn
    hip.download()
    sec = time.perf_counter() - tic
    return {"workload": f"RADAU tol {tol:g} with 1PN + radiation pressure/PR drag (beta 1e-4) + Yarkovsky "
                        f"(1e-12 AU/day^2), 8 planets + {nparticles} particles, {days:g} d, first step 1 d",
            "seconds": round(sec, 3), "sequences": nseq, "force_evaluations": forces,
            "particle_steps_per_s": round(nparticles * nseq / sec, 1),
            "reference_published": PUBLISHED_EXTRA_FORCES,
            "speedup_vs_reference_cuda": round(PUBLISHED_EXTRA_FORCES["cuda_rtx4070_s"] / sec, 2)}


def _cpu_ra15_days_worker(job):
    sub, days, tol = job

    def run():
        from oracle.bindings import Reference
        R = Reference(nmax=len(sub["m"]) + 16, cmax=len(sub["m"]) + 16)
        R.ra15_run(sub, 1, pn=1, tol=tol)              # warm
        t = time.perf_counter()
        _, _, hd, _ = R.ra15_run(sub, 0, pn=1, tol=tol, until=days)
        return time.perf_counter() - t, len(hd)
    return _in_big_thread(run)


def cpu_radau_extra_forces(nparticles=100_000, days=365.0, tol=1e-11, nproc=16):
    """The reference mdt_ra15 with mfo_all (1PN, mfo_pr, mfo_ngf) on the same
    case split into `nproc` disjoint particle shards, each integrated for the
    full 365 days in its own process: the whole case's wall time on nproc
    cores (each shard takes its own step sizes)."""
    import multiprocessing as mp
    import numpy as np
    from mercuda_amd import systems
    from oracle.bindings import REF_PATH
    if not REF_PATH.exists():
        return None
    s = systems.config_extra_forces(nparticles=nparticles)
    nb = s["nbig"]
    shard = nparticles // nproc
    jobs = []
    for p in range(nproc):
        idx = np.concatenate([np.arange(nb), nb + p * shard + np.arange(shard)])
        sub = {k: s[k][idx] for k in ("m", "x", "v", "rce", "rphys", "ngf")}
        sub.update(nbig=nb, h=s["h"], rcen=s["rcen"], jcen=np.zeros(3), ngflag=s["ngflag"])
        jobs.append((sub, days, tol))
    with mp.get_context("spawn").Pool(nproc) as pool:
        res = pool.map(_cpu_ra15_days_worker, jobs)
    return max(r[0] for r in res), shard, [r[1] for r in res]


def bs_rate(hip, algor, nparticles=1_000_000, nsteps=8):
    """Algorithm 2 (mdt_bs1) or 3 (mdt_bs2) with mfo_all, tol 1e-12, on the
    MVS workload's 8 planets + 1M particles, first step 4 d: accepted steps x
    particles per second through mercury_hip_step (one host round trip per
    attempt for the step control), after 2 warm-up steps."""
    from mercuda_amd import systems
    s = systems.config_large(seed=2, nparticles=nparticles)
    hip.configure(algor)
    hip.upload_system(s)
    h, t = s["h"], 0.0
    for _ in range(2):
        h, hd, _, _ = hip.step(t, h, 1e-12)
        t += hd
    days = 0.0
    rej = 0
    tic = time.perf_counter()
    for _ in range(nsteps):
        h, hd, _, r = hip.step(t, h, 1e-12)
        t += hd
        days += hd
        rej += r
    sec = time.perf_counter() - tic
    return {"workload": f"{'BS' if algor == 2 else 'BS2'} (algorithm {algor}) tol 1e-12, 8 planets + "
                        f"{nparticles} particles", "steps": nsteps,
            "particle_steps_per_s": round(nparticles * nsteps / sec, 1),
            "ms_per_step": round(sec * 1e3 / nsteps, 3), "mean_step_days": round(days / nsteps, 4),
            "halvings": rej}


def _cpu_bs_worker(job):
    sub, algor, nsteps = job

    def run():
        from oracle.bindings import Reference
        R = Reference(nmax=len(sub["m"]) + 16, cmax=len(sub["m"]) + 16)
        R.bs_run(algor, sub, 1)                         # warm
        t = time.perf_counter()
        R.bs_run(algor, sub, nsteps)
        return time.perf_counter() - t
    return _in_big_thread(run)


def cpu_bs(sysd, algor, nproc=16, shard=62500, nsteps=8):
    """The reference mdt_bs1 / mdt_bs2 with mfo_all on `nproc` disjoint particle
    shards of the 1M system in parallel processes (each shard takes its own
    step sizes, as the error control spans the bodies it integrates).  Returns
    particle-steps/s."""
    import multiprocessing as mp
    import numpy as np
    from oracle.bindings import REF_PATH
    if not REF_PATH.exists():
        return None
    nb = sysd["nbig"]
    jobs = []
    for p in range(nproc):
        idx = np.concatenate([np.arange(nb), nb + p * shard + np.arange(shard)])
        sub = {k: sysd[k][idx] for k in ("m", "x", "v", "rce", "rphys")}
        sub.update(nbig=nb, h=sysd["h"], rcen=sysd["rcen"], jcen=np.zeros(3))
        jobs.append((sub, algor, nsteps))
    with mp.get_context("spawn").Pool(nproc) as pool:
        times = pool.map(_cpu_bs_worker, jobs)
    return nproc * shard * nsteps / max(times), max(times)


def hybrid_rate(hip, nparticles=200_000, nsteps=3, driver_path=False):
    """HYBRID on 8 planets + 200k particles, half within 0.1 AU of a planet's
    orbit (changeover 3 Hill radii, dt 2 d): every step's device work --
    regular map, changeover scan, encounter members' BS2 substeps through the
    compact context -- driven by mercuda_amd.hybrid as the driver drives it.
    driver_path False: the encounter loop's step control on the device
    (mercury_hip_encounter_advance, one queued launch per BS2 attempt).
    driver_path True: the path the unchanged Fortran driver takes, one
    mercury_hip_step per BS2 substep (the step control on the host) and the
    members' states downloaded after each substep for the driver's encounter
    listing.  The listing itself (the driver's own mce_stat on the host) and
    its output are not in the timed loop."""
    from mercuda_amd import hybrid, systems
    s = systems.config_encounters(seed=3, nparticles=nparticles)
    hip.configure(10)
    hip.upload_system(s)
    hybrid.hy_run(hip, s, 1, upload=False)
    hip.configure(10)
    hip.upload_system(s)
    hybrid.mdt_hkce.substeps = 0
    # the driver's path: each substep's states come back to the host
    pull = (lambda *a: None) if driver_path else None
    tic = time.perf_counter()
    _, _, most = hybrid.hy_run(hip, s, nsteps, upload=False, on_substep=pull,
                               progress=lambda k, npairs: log(f"hybrid step {k}: {npairs} encounter pairs, "
                                                              f"{hybrid.mdt_hkce.substeps} BS2 substeps so far"))
    sec = time.perf_counter() - tic
    return {"workload": f"HYBRID, 8 planets + {nparticles} particles (half near planet orbits), dt 2 d",
            "encounter_loop": "per-substep mercury_hip_step + state download (the driver's path)" if driver_path
            else "device loop (mercury_hip_encounter_advance)",
            "steps": nsteps, "particle_steps_per_s": round(nparticles * nsteps / sec, 1),
            "ms_per_step": round(sec * 1e3 / nsteps, 3), "most_encounter_pairs_in_a_step": most,
            "bs2_substeps": hybrid.mdt_hkce.substeps}


def copy_probe(mib):
    """A plain device-to-device copy of known size (three hipMemcpy calls of
    `mib` MiB), the calibration run for rocprofv3's HBM byte counters
    (FETCH_SIZE, WRITE_SIZE): the runtime's copy kernel reads and writes
    exactly `mib` MiB per dispatch."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3) != 0:      # hipMemcpyDeviceToDevice
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copied {mib} MiB device-to-device 3 times")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2048)
    ap.add_argument("--warmup", type=int, default=1024)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--spl", type=int, default=32, help="steps fused per particle launch")
    ap.add_argument("--block", type=int, default=128)
    ap.add_argument("--ppl", type=int, default=1, help="particles per lane of the fast particle kernel")
    ap.add_argument("--no-screen", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=16)
    ap.add_argument("--no-extras", action="store_true", help="skip the ensemble/single ratio")
    ap.add_argument("--extras", default="all",
                    help="comma list of extra workloads to run: ensemble, adjoint, radau, radau_extra_forces, "
                         "bs2, bs3, hybrid, hybrid_driver_path, hybrid_20k, hybrid_20k_driver_path (default all)")
    ap.add_argument("--autotune", action="store_true",
                    help="let the autotuner pick steps per launch and workgroup size instead of --spl/--block "
                         "(the defaults are its usual choice on MI355X)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the HBM-counter calibration copy of MIB MiB and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the massless-particle path has no multi-GPU mode in this build "
              "(particles would shard across GPUs with a replicated planet table).",
              file=sys.stderr)
        sys.exit(2)

    from mercuda_amd import systems
    from mercuda_amd.hip import MercuryHip

    hip = MercuryHip()
    if not hip.available():
        raise SystemExit("no HIP device")
    log(f"generating 8 planets + {args.particles} particles")
    sysd = systems.config_large(seed=2, nparticles=args.particles)
    hip.configure(1)
    hip.upload_system(sysd)
    hip.set_tuning(steps_per_launch=args.spl, block_size=args.block, particles_per_lane=args.ppl)
    screen = not args.no_screen
    h = sysd["h"]
    spl, block, ppl = args.spl, args.block, args.ppl
    if args.autotune:
        log("autotuning steps per launch, workgroup size and particles per lane (128-step probes, state restored)")
        tune = hip.autotune(h=h, probe_steps=128)
        spl, block, ppl = tune["steps_per_launch"], tune["block"], tune["particles_per_lane"]
        log(f"autotuner: {tune}")
    log(f"warmup {args.warmup} steps")
    hip.advance(0.0, h, args.warmup, sysd["rcen"], screen=screen)
    t0 = args.warmup * h
    # HIP events around the fast kernel's launches only: events on the planet
    # and slow-set streams would add their own latency to the timed call
    hip.timing(True, fast_only=True)
    log(f"timed {args.steps} steps")
    tic = time.perf_counter()
    hip.advance(t0, h, args.steps, sysd["rcen"], screen=screen)
    sec = time.perf_counter() - tic
    tm = hip.timing_read()
    fast_steps, gen_steps = hip.work_read()
    nevents = len(hip.advance_events())
    hip.timing(False)
    # every kernel kind timed on a second call of the same length, after the timed one
    hip.timing(True)
    hip.advance(t0 + args.steps * h, h, args.steps, sysd["rcen"], screen=screen)
    tm2 = hip.timing_read()
    hip.timing(False)
    npart = args.particles
    value = npart * args.steps / sec
    log(f"{value:.4e} particle-steps/s; kernels {tm}; events {nevents}; parked {hip.parked()}")

    # roofline of the dominant kernel (fast particle kernel)
    launches = max(tm["particle_launches"], 1)
    avg_s = tm["particle_ms"] / launches / 1e3
    flop_per_launch = FLOP_PER_PSTEP * fast_steps / launches
    achieved = flop_per_launch / avg_s / 1e12
    roofline = {
        # compute roof: the FP64 vector peak (78.6 TFLOP/s on MI355X); the
        # kernel runs on the FP64 VALU -- nothing here is a dense contraction
        "bound": "fp64_valu",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
        "traffic": round(TRAFFIC_PER_PARTICLE_LAUNCH * npart),
        "kernel": "mvs_particles_fast",
        "avg_launch_ms": round(avg_s * 1e3, 4),
        "flop_per_particle_step": FLOP_PER_PSTEP,
        "particle_steps_per_launch": round(fast_steps / launches, 1),
        "hbm_gbs_algorithmic": round(BYTES_PER_PARTICLE_LAUNCH * npart / avg_s / 1e9, 1),
        "hbm_bytes_algorithmic_per_launch": round(BYTES_PER_PARTICLE_LAUNCH * npart),
        "time_share": {"fast_ms": round(tm["particle_ms"], 2), "wall_ms": round(sec * 1e3, 2),
                       "second_call": {"fast_ms": round(tm2["particle_ms"], 2),
                                       "general_ms": round(tm2["general_ms"], 2),
                                       "planet_ms": round(tm2["planet_ms"], 2)}},
    }
    extras = {}
    names = ("ensemble", "adjoint", "radau", "radau_extra_forces", "bs2", "bs3", "hybrid", "hybrid_driver_path",
             "hybrid_20k", "hybrid_20k_driver_path")
    want = set() if args.no_extras else (set(names) if args.extras == "all" else set(args.extras.split(",")))
    if want - set(names):
        raise SystemExit(f"unknown extras {sorted(want - set(names))}; choose from {names}")
    if "ensemble" in want:
        extras["ensemble"] = ensemble_ratio(hip, steps=128, spl=spl, block=block)
    if "adjoint" in want:
        log("adjoint: orbit-fitting workload")
        extras["adjoint"] = adjoint_cost(hip)
    if "radau" in want:
        log("radau: 8 planets + 1M particles")
        extras["radau"] = radau_rate(hip)
    if "radau_extra_forces" in want:
        log("radau with 1PN + PR + Yarkovsky: 8 planets + 100k particles, 365 d")
        extras["radau_extra_forces"] = radau_extra_forces(hip)
    for algor in (2, 3):
        if f"bs{algor}" in want:
            log(f"bs: algorithm {algor}, 8 planets + 1M particles")
            extras[f"bs{algor}"] = bs_rate(hip, algor)
    if "hybrid" in want:
        log("hybrid: 8 planets + 200k particles")
        extras["hybrid"] = hybrid_rate(hip)
    if "hybrid_driver_path" in want:
        log("hybrid, the driver's per-substep path: 8 planets + 200k particles")
        extras["hybrid_driver_path"] = hybrid_rate(hip, nsteps=1, driver_path=True)
    if "hybrid_20k" in want:
        log("hybrid: 8 planets + 20k particles")
        extras["hybrid_20k"] = hybrid_rate(hip, nparticles=20_000, nsteps=5)
    if "hybrid_20k_driver_path" in want:
        log("hybrid, the driver's per-substep path: 8 planets + 20k particles")
        extras["hybrid_20k_driver_path"] = hybrid_rate(hip, nparticles=20_000, nsteps=5, driver_path=True)
    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline: reference mdt_mvs on {args.cpu_procs} processes")
        r = cpu_baseline(sysd, nproc=args.cpu_procs)
        if r:
            rate, tmax = r
            cpu = {"value": round(rate, 1), "unit": "particle-steps/s", "cores": args.cpu_procs,
                   "kind": "reference",
                   "sample": f"{args.cpu_procs} processes x 31250 particles x 1536 steps of this workload "
                             f"(reference mdt_mvs + mce_cent, flang -O3 -ffp-contract=off), "
                             f"slowest process {tmax:.2f} s",
                   "gpu_over_cpu": round(value / rate, 1)}
        if extras.get("radau"):
            log(f"cpu baseline: reference mdt_ra15 on {args.cpu_procs} processes")
            r = cpu_radau(sysd, nproc=args.cpu_procs)
            if r:
                extras["radau"]["cpu_baseline"] = {
                    "value": round(r[0], 1), "unit": "particle-steps/s", "cores": args.cpu_procs, "kind": "reference",
                    "sample": f"{args.cpu_procs} processes x 62500 particles x 8 RA15 sequences of this system "
                              f"(tol 1e-12, 1PN; each shard takes its own step sizes), slowest process {r[1]:.2f} s",
                    "gpu_over_cpu": round(extras["radau"]["particle_steps_per_s"] / r[0], 1)}
        if extras.get("radau_extra_forces"):
            log(f"cpu baseline: reference mdt_ra15 with 1PN + PR + Yarkovsky, 365 d, {args.cpu_procs} processes")
            r = cpu_radau_extra_forces(nproc=args.cpu_procs)
            if r:
                ex = extras["radau_extra_forces"]
                ex["cpu_baseline"] = {
                    "value": round(r[0], 3), "unit": "s", "cores": args.cpu_procs, "kind": "reference",
                    "sample": f"the whole case: {args.cpu_procs} processes x {r[1]} particles, each shard "
                              f"365 d with its own step sizes ({min(r[2])}-{max(r[2])} sequences); wall time of "
                              f"the slowest",
                    "cpu_over_gpu_time": round(r[0] / ex["seconds"], 1)}
        for algor in (2, 3):
            if extras.get(f"bs{algor}"):
                log(f"cpu baseline: reference algorithm {algor} on {args.cpu_procs} processes")
                r = cpu_bs(sysd, algor, nproc=args.cpu_procs)
                if r:
                    extras[f"bs{algor}"]["cpu_baseline"] = {
                        "value": round(r[0], 1), "unit": "particle-steps/s", "cores": args.cpu_procs,
                        "kind": "reference",
                        "sample": f"{args.cpu_procs} processes x 62500 particles x 8 steps of this system "
                                  f"(mdt_bs{'1' if algor == 2 else '2'} with mfo_all, tol 1e-12; each shard takes "
                                  f"its own step sizes), slowest process {r[1]:.2f} s",
                        "gpu_over_cpu": round(extras[f"bs{algor}"]["particle_steps_per_s"] / r[0], 1)}
        if extras.get("hybrid_20k"):
            log("cpu baseline: reference mdt_hy on 20k particles, one process")
            r = cpu_hybrid()
            if r:
                extras["hybrid_20k"]["cpu_baseline"] = {
                    "value": round(r[0], 1), "unit": "particle-steps/s", "cores": 1, "kind": "reference",
                    "sample": f"1 HYBRID step of the same 20k-particle system (the reference integrates a system "
                              f"serially), {r[1]:.2f} s",
                    "gpu_over_cpu": round(extras["hybrid_20k"]["particle_steps_per_s"] / r[0], 1)}
    line = {
        "metric": "FP64 particle-steps/s, MVS 8pl+1M TP; batched ensemble & adjoint/fwd cost",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"MVS, 8 planets + {npart} massless particles, dt 4 d, "
                               f"encounter screening {'on' if screen else 'off'}",
                   "particles": npart, "planets": 8, "dt_days": h,
                   "steps_per_launch": spl, "block": block, "particles_per_lane": ppl,
                   "tuning": "autotuned" if args.autotune else "fixed", "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
