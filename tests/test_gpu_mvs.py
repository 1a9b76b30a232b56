"""Parity of the HIP MVS path with the oracle (C restatement of mdt_mvs,
mce_stat and mce_cent) on identical seeded inputs.

Reference behaviour being matched: the MVS step mdt_mvs
(mercury6_2.for:3987-4097) with the driver's central-impact check
(mal_hcon mercury6_2.for:811-841); parity bar as the reference's own CPU-vs-GPU
test (tests/test_gpu_algorithms.py:15-32), tightened to 1e-11 relative.
"""
import numpy as np
import pytest

from helpers import REL_TOL, adv_event_rows, compare_rows, hip_event_rows, oracle_rows, rel_err
from mercuda_amd import systems

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def run_stepwise(hip, sysd, nsteps, time0=0.0):
    """The drop-in path: one mercury_hip_step + mercury_hip_events per step."""
    hip.configure(1)
    hip.upload_system(sysd)
    h = sysd["h"]
    time = time0
    rows = []
    for s in range(nsteps):
        hip.step(time, h)
        time = time + h
        rows += hip_event_rows(hip.events(time, h, sysd["rcen"]), s)
    x, v = hip.download()
    return x, v, rows


def run_fused(hip, sysd, nsteps, spl=32, time0=0.0):
    hip.configure(1)
    hip.upload_system(sysd)
    hip.set_tuning(steps_per_launch=spl)
    hip.advance(time0, sysd["h"], nsteps, sysd["rcen"], screen=True)
    x, v = hip.download()
    return x, v, adv_event_rows(hip.advance_events())


def check_state(x, v, xo, vo, label):
    ex, nx = rel_err(x[1:], xo[1:])
    ev, nv = rel_err(v[1:], vo[1:])
    say(f"{label}: max rel err x {ex:.3e} v {ev:.3e}; bodies differing {nx}/{len(x) - 1}")
    assert ex <= REL_TOL and ev <= REL_TOL


def test_stepwise_matches_oracle(hip, oracle):
    sysd = systems.config_small(seed=1, nparticles=1000)
    say("case: stepwise 8 planets + 1000 particles, 456 steps of 8 d")
    x, v, rows = run_stepwise(hip, sysd, 456)
    xo, vo, evi, evd = oracle.mvs_run(sysd, 456)
    check_state(x, v, xo, vo, "stepwise")
    compare_rows(rows, oracle_rows(evi, evd))


@pytest.mark.parametrize("spl", [1, 7, 32])
def test_fused_matches_oracle(hip, oracle, spl):
    sysd = systems.config_small(seed=1, nparticles=1000)
    say(f"case: fused advance S={spl}, 8 planets + 1000 particles, 456 steps")
    x, v, rows = run_fused(hip, sysd, 456, spl=spl)
    xo, vo, evi, evd = oracle.mvs_run(sysd, 456)
    check_state(x, v, xo, vo, f"fused S={spl}")
    compare_rows(rows, oracle_rows(evi, evd))


def test_fused_equals_stepwise_bitwise(hip):
    sysd = systems.config_large(seed=5, nparticles=5000)
    say("case: fused vs stepwise bitwise, 5000 particles, 64 steps")
    x1, v1, r1 = run_stepwise(hip, sysd, 64)
    x2, v2, r2 = run_fused(hip, sysd, 64, spl=16)
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2)
    compare_rows(r2, r1, rtol=0.0)


@pytest.mark.parametrize("count", [0, 1, 63, 64, 65, 255, 256, 257, 4095, 4096, 4097])
def test_ragged_particle_counts(hip, oracle, count):
    sysd = systems.config_small(seed=11, nparticles=count)
    say(f"case: ragged count {count}, 40 steps")
    x, v, rows = run_fused(hip, sysd, 40, spl=16)
    xo, vo, evi, evd = oracle.mvs_run(sysd, 40)
    check_state(x, v, xo, vo, f"count {count}")
    compare_rows(rows, oracle_rows(evi, evd))


def test_no_planets_pure_kepler(hip, oracle):
    sysd = systems.make_system(21, 3000, nplanets=0, a_range=(0.2, 30.0), emax=0.9, h=8.0)
    say("case: central body only, 3000 particles e<0.9, 200 steps")
    x, v, rows = run_fused(hip, sysd, 200, spl=25)
    xo, vo, evi, evd = oracle.mvs_run(sysd, 200)
    check_state(x, v, xo, vo, "no planets")
    compare_rows(rows, oracle_rows(evi, evd))


def test_kepler_regimes(hip, oracle):
    """Eccentric, near-parabolic and hyperbolic particles: every branch of
    drift_dan/drift_kepu (kepmd, Newton, Laguerre, p3solve) through the
    general kernel, at 0.7/1.0/1.2 x escape speed as the reference's
    tests/test_gpu_algorithms.py:67-77."""
    rng = np.random.default_rng(7)
    sysd = systems.config_small(seed=2, nparticles=0)
    mu = sysd["m"][0]
    k = 3000
    r = 10 ** rng.uniform(-0.5, 1.2, k)
    vesc = np.sqrt(2 * mu / r)
    f = rng.choice([0.3, 0.7, 0.99, 1.0, 1.01, 1.2, 2.0], k)
    dirx = rng.normal(size=(k, 3)); dirx /= np.linalg.norm(dirx, axis=1)[:, None]
    dirv = rng.normal(size=(k, 3)); dirv /= np.linalg.norm(dirv, axis=1)[:, None]
    xp = dirx * r[:, None]; vp = dirv * (f * vesc)[:, None]
    sysd["m"] = np.concatenate([sysd["m"], np.zeros(k)])
    sysd["x"] = np.concatenate([sysd["x"], xp]); sysd["v"] = np.concatenate([sysd["v"], vp])
    systems.finish_limits(sysd)
    branches = np.bincount([oracle.drift_branch(mu, np.concatenate([xp[q], vp[q]]), sysd["h"])
                            for q in range(k)], minlength=4)
    say(f"case: kepler regimes, branch counts kepmd/newton/laguerre/fail = {branches.tolist()}")
    x, v, rows = run_fused(hip, sysd, 100, spl=10)
    xo, vo, evi, evd = oracle.mvs_run(sysd, 100)
    check_state(x, v, xo, vo, "kepler regimes")
    compare_rows(rows, oracle_rows(evi, evd))


def regimes_system(seed, k=3000):
    """Particles at 0.3-2x the local escape speed: many leave the kepmd branch
    (slow-set members of the general kernel), some for a few steps only."""
    rng = np.random.default_rng(seed)
    sysd = systems.config_small(seed=2, nparticles=0)
    mu = sysd["m"][0]
    r = 10 ** rng.uniform(-0.5, 1.2, k)
    vesc = np.sqrt(2 * mu / r)
    f = rng.choice([0.3, 0.7, 0.9, 0.99, 1.0, 1.2], k)
    dirx = rng.normal(size=(k, 3)); dirx /= np.linalg.norm(dirx, axis=1)[:, None]
    dirv = rng.normal(size=(k, 3)); dirv /= np.linalg.norm(dirv, axis=1)[:, None]
    sysd["m"] = np.concatenate([sysd["m"], np.zeros(k)])
    sysd["x"] = np.concatenate([sysd["x"], dirx * r[:, None]])
    sysd["v"] = np.concatenate([sysd["v"], dirv * (f * vesc)[:, None]])
    systems.finish_limits(sysd)
    return sysd


def test_split_calls_equal_one_call_bitwise(hip):
    """The slow set persists from call to call (members kept, members leaving
    when their drifts return to the kepmd branch); calls of any length, and a
    reset of the acceleration history between them, give bitwise the states
    and events of one long call."""
    sysd = regimes_system(17)
    h = sysd["h"]
    say("case: split calls 1+5+20+3+37 (reset before the 4th) vs one call of 66 steps")
    hip.configure(1)
    hip.upload_system(sysd)
    hip.set_tuning(steps_per_launch=16)
    hip.advance(0.0, h, 66, sysd["rcen"], screen=True)
    x1, v1 = hip.download()
    r1 = adv_event_rows(hip.advance_events())
    hip.upload_system(sysd)
    t, done, r2, members = 0.0, 0, [], []
    for k, n in enumerate([1, 5, 20, 3, 37]):
        if k == 3:
            hip.reset(0)
        hip.advance(t, h, n, sysd["rcen"], screen=True)
        members.append(hip.parked())
        r2 += [(row[0] + done,) + tuple(row[1:]) for row in adv_event_rows(hip.advance_events())]
        for _ in range(n):
            t = t + h
        done += n
    x2, v2 = hip.download()
    say(f"slow-set members per call {members}")
    assert min(members) > 0
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2)
    compare_rows(sorted(r2), sorted(r1), rtol=0.0)


def test_encounters_fire(hip, oracle):
    """Half the particles near planet orbits: encounter minima, near misses and
    central impacts are listed identically (indices, kind, step)."""
    sysd = systems.config_encounters(seed=3, nparticles=4000)
    sysd["rce"][1:sysd["nbig"]] *= 3.0
    sysd["rcen"] = 0.05
    say("case: encounter-rich 4000 particles, 300 steps of 2 d")
    x, v, rows = run_fused(hip, sysd, 300, spl=32)
    xo, vo, evi, evd = oracle.mvs_run(sysd, 300)
    kinds = np.bincount(evi[:, 1], minlength=9) if len(evi) else []
    say(f"oracle events {len(evi)} by kind {list(kinds)}")
    assert len(evi) > 50
    check_state(x, v, xo, vo, "encounters")
    compare_rows(rows, oracle_rows(evi, evd))
    x2, v2, rows2 = run_stepwise(hip, sysd, 300)
    compare_rows(rows2, oracle_rows(evi, evd))


def test_long_run_1e4_steps(hip, oracle):
    """The parity gate: per-particle x, v within 1e-11 after 10^4 MVS steps."""
    sysd = systems.config_small(seed=1, nparticles=1000)
    say("case: 10^4 steps, 8 planets + 1000 particles")
    x, v, rows = run_fused(hip, sysd, 10000, spl=64)
    xo, vo, evi, evd = oracle.mvs_run(sysd, 10000)
    check_state(x, v, xo, vo, "1e4 steps")
    compare_rows(rows, oracle_rows(evi, evd))


def test_full_size_sampled(hip, oracle):
    """1M particles on the GPU; particles never interact, so a random sample of
    them run alone through the oracle must give the same states."""
    sysd = systems.config_large(seed=2, nparticles=1_000_000)
    say("case: 1M particles, 48 steps of 4 d, sampled parity")
    x, v, rows = run_fused(hip, sysd, 48, spl=16)
    rng = np.random.default_rng(0)
    nb = sysd["nbig"]
    pick = np.sort(rng.choice(np.arange(nb, len(sysd["m"])), 3000, replace=False))
    keep = np.concatenate([np.arange(nb), pick])
    sub = dict(sysd)
    for k in ("m", "x", "v", "rce", "rphys", "rcrit"):
        sub[k] = sysd[k][keep]
    xo, vo, evi, evd = oracle.mvs_run(sub, 48)
    check_state(x[keep], v[keep], xo, vo, "1M sampled")


def test_export_corrector(hip, oracle):
    """export(physical=1) applies the symplectic corrector mco_mvs2h
    (mercury6_2.for:2107-2219) to a copy; the state itself is untouched."""
    sysd = systems.config_small(seed=1, nparticles=500)
    say("case: corrector export after 50 steps")
    x, v, _ = run_stepwise(hip, sysd, 50)
    xc, vc = hip.export(sysd["h"], physical=True)
    xo, vo, _, _ = oracle.mvs_run(sysd, 50)
    xh, vh = oracle.mvs2h(sysd, xo, vo)
    check_state(xc, vc, xh, vh, "corrector")
    x2, v2 = hip.download()
    assert np.array_equal(x2, x) and np.array_equal(v2, v)
    xp, vp = hip.export(sysd["h"], physical=False)
    assert np.array_equal(xp, x) and np.array_equal(vp, v)
