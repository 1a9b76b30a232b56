"""Batched ensembles: many independent systems of uneven size (2..12 planets,
log-uniform particle counts) advanced together must give, system by system,
the states and event lists of the oracle run on each system alone."""
import numpy as np
import pytest

from helpers import REL_TOL, compare_rows, hip_event_rows, oracle_rows, rel_err
from mercuda_amd import systems

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rows_by_system(events, nsys):
    out = [[] for _ in range(nsys)]
    per = {}
    for ev in events:
        per.setdefault((ev[0], ev[1]), []).append(ev[2:])
    for (sysid, step) in sorted(per):
        out[sysid] += hip_event_rows(per[(sysid, step)], step)
    return out


def test_ensemble_matches_oracle(hip, oracle):
    ens = systems.ensemble(seed=9, nsystems=24, nmin=20, nmax=3000)
    for q in ens[::3]:
        q["rce"][1:q["nbig"]] *= 4.0        # more encounter events
    nsteps = 120
    say(f"case: ensemble of {len(ens)} systems, {sum(len(q['m']) for q in ens)} bodies, {nsteps} steps")
    hip.configure(1)
    hip.ensemble_upload(ens)
    hip.set_tuning(steps_per_launch=32)
    hip.ensemble_advance(0.0, nsteps, screen=True)
    states = hip.ensemble_download()
    evs = rows_by_system(hip.advance_events(), len(ens))
    for k, q in enumerate(ens):
        xo, vo, evi, evd = oracle.mvs_run(q, nsteps)
        x, v = states[k]
        ex, _ = rel_err(x[1:], xo[1:])
        ev_, _ = rel_err(v[1:], vo[1:])
        assert ex <= REL_TOL and ev_ <= REL_TOL, (k, ex, ev_)
        compare_rows(evs[k], oracle_rows(evi, evd))
    say(f"events per system: {[len(e) for e in evs]}")


def test_ensemble_equals_single_runs_bitwise(hip):
    ens = systems.ensemble(seed=10, nsystems=6, nmin=100, nmax=2000)
    nsteps = 70
    say(f"case: ensemble of {len(ens)} vs single-system runs, {nsteps} steps")
    hip.configure(1)
    hip.ensemble_upload(ens)
    hip.set_tuning(steps_per_launch=16)
    hip.ensemble_advance(0.0, nsteps, screen=True)
    states = hip.ensemble_download()
    for k, q in enumerate(ens):
        hip.configure(1)
        hip.upload_system(q)
        hip.advance(0.0, q["h"], nsteps, q["rcen"], screen=True)
        x, v = hip.download()
        assert np.array_equal(x, states[k][0]) and np.array_equal(v, states[k][1]), k


def test_ensemble_of_one_system(hip, oracle):
    """An ensemble may hold a single system: it keeps its own step and
    central radius (the advance call passes none)."""
    ens = systems.ensemble(seed=14, nsystems=1, nmin=500, nmax=3000)
    q = ens[0]
    q["rce"][1:q["nbig"]] *= 4.0
    nsteps = 90
    say(f"case: ensemble of one system, {len(q['m'])} bodies, {nsteps} steps")
    hip.configure(1)
    hip.ensemble_upload(ens)
    hip.set_tuning(steps_per_launch=16)
    hip.ensemble_advance(0.0, 40, screen=True)
    hip.ensemble_advance(40 * q["h"], nsteps - 40, screen=True)
    (x, v), = hip.ensemble_download()
    xo, vo, evi, evd = oracle.mvs_run(q, nsteps)
    ex, _ = rel_err(x[1:], xo[1:])
    ev_, _ = rel_err(v[1:], vo[1:])
    assert ex <= REL_TOL and ev_ <= REL_TOL, (ex, ev_)


def test_particles_per_lane_bitwise(hip):
    """Lanes that take 2 or 4 particles of their tile one after the other give
    the states and event lists of one particle per lane, bit for bit: on one
    system whose particle count fills no whole tile, and on an ensemble."""
    one = systems.config_small(seed=21, nparticles=1000)
    one["rce"][1:one["nbig"]] *= 4.0
    ens = systems.ensemble(seed=22, nsystems=7, nmin=50, nmax=1500)
    nsteps = 48

    def run_one(ppl):
        hip.configure(1)
        hip.upload_system(one)
        hip.set_tuning(steps_per_launch=16, block_size=128, particles_per_lane=ppl)
        hip.advance(0.0, one["h"], nsteps, one["rcen"], screen=True)
        x, v = hip.download()
        return x, v, sorted(hip.advance_events())

    def run_ens(ppl):
        hip.configure(1)
        hip.ensemble_upload(ens)
        hip.set_tuning(steps_per_launch=16, block_size=64, particles_per_lane=ppl)
        hip.ensemble_advance(0.0, nsteps, screen=True)
        return hip.ensemble_download(), sorted(hip.advance_events())

    try:
        base_one = run_one(1)
        base_ens = run_ens(1)
        for ppl in (2, 4):
            say(f"case: particles per lane {ppl}, one system of 1000 particles and an ensemble of {len(ens)}")
            x, v, ev = run_one(ppl)
            assert np.array_equal(x, base_one[0]) and np.array_equal(v, base_one[1]), ppl
            assert ev == base_one[2], ppl
            states, evs = run_ens(ppl)
            for k in range(len(ens)):
                assert np.array_equal(states[k][0], base_ens[0][k][0]), (ppl, k)
                assert np.array_equal(states[k][1], base_ens[0][k][1]), (ppl, k)
            assert evs == base_ens[1], ppl
    finally:
        hip.set_tuning(steps_per_launch=32, block_size=128, particles_per_lane=1)
