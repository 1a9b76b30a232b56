"""Device branches the other parity tests do not reach, and the full-size
configurations, against the oracle:

- negative steps (backward integration) for MVS (fused and per step), RADAU,
  BS, BS2 and HYBRID (reference tests/test_gpu_algorithms.py:43-47,96-100);
- central-body oblateness (J2, J4, J6) on the MVS and HYBRID device paths
  (tests/test_gpu_algorithms.py:79-87);
- the event list growing past its capacity within one step's screen
  (tests/test_events.py:19-53), and an advance call whose events overflow
  failing loudly;
- RA15 on 100,000 particles of the MVS/RADAU workload's generator (tol 1e-12,
  1PN): the step-size law takes max |b7| over every body, so the sizes must
  be the oracle's bit for bit;
- the 512-system ensemble, sampled systems against the oracle run alone;
- HYBRID on 50,000 particles of the HYBRID workload's generator (a quarter of
  its 200,000: the oracle's single-threaded BS2 over the 200,000-particle
  system's ~25,000-pair encounter set takes more than 15 minutes a step),
  with the encounter loop's step control on the device."""
import numpy as np
import pytest

from helpers import REL_TOL, adv_event_rows, compare_rows, hip_event_rows, oracle_rows, rel_err
from hybrid_driver import hy_run_device
from mercuda_amd import systems

pytestmark = pytest.mark.gpu

JC = (1e-6 * 0.0046 ** 2, -2e-7 * 0.0046 ** 4, 5e-8 * 0.0046 ** 6)


def say(*a):
    print(*a, flush=True)


def check(x, v, xo, vo, label, tol=REL_TOL):
    ex, nx = rel_err(x[1:], xo[1:])
    ev, nv = rel_err(v[1:], vo[1:])
    say(f"{label}: max rel err x {ex:.3e} v {ev:.3e}; bodies differing {nx}/{len(x) - 1}")
    assert ex <= tol and ev <= tol


def mvs_fused(hip, s, nsteps, spl=16):
    hip.configure(1)
    hip.upload_system(s)
    hip.set_tuning(steps_per_launch=spl)
    hip.advance(0.0, s["h"], nsteps, s["rcen"], screen=True)
    x, v = hip.download()
    return x, v, adv_event_rows(hip.advance_events())


def mvs_stepwise(hip, s, nsteps):
    hip.configure(1)
    hip.upload_system(s)
    t, rows = 0.0, []
    for k in range(nsteps):
        hip.step(t, s["h"])
        t = t + s["h"]
        rows += hip_event_rows(hip.events(t, s["h"], s["rcen"]), k)
    x, v = hip.download()
    return x, v, rows


def encounter_system(seed, n, h):
    s = systems.make_system(seed, n, a_range=(0.5, 40.0), emax=0.3, imax_deg=20.0, h=h, near_fraction=0.5)
    s["rce"][1:s["nbig"]] *= 3.0
    s["rcen"] = 0.05
    s["jcen"] = np.zeros(3)
    return s


@pytest.mark.parametrize("jcen", [None, JC])
def test_mvs_backward_and_oblate(hip, oracle, jcen):
    s = encounter_system(41, 2000, -4.0 if jcen is None else 4.0)
    if jcen is not None:
        s["jcen"] = np.array(jcen)
    label = "h < 0" if jcen is None else "J2/J4/J6"
    say(f"case: MVS {label}, 8 planets + 2000 particles, 200 steps, fused and per step")
    xo, vo, evi, evd = oracle.mvs_run(s, 200)
    kinds = sorted(set(int(k) for k in evi[:, 1])) if len(evi) else []
    say(f"oracle events {len(evi)}, kinds {kinds}")
    assert len(evi) > 10
    x, v, rows = mvs_fused(hip, s, 200)
    check(x, v, xo, vo, f"MVS fused {label}")
    compare_rows(rows, oracle_rows(evi, evd))
    x, v, rows = mvs_stepwise(hip, s, 200)
    check(x, v, xo, vo, f"MVS per step {label}")
    compare_rows(rows, oracle_rows(evi, evd))


def test_radau_backward(hip, oracle):
    s = systems.config_small(seed=12, nparticles=800)
    say("case: RA15 backward (first trial -4 d), 1PN, 8 planets + 800 particles, 120 sequences")
    hip.configure(4)
    hip.upload_system(s, pn=1)
    h, t, hd = -4.0, 0.0, []
    for _ in range(120):
        h, hdid, nf, nr = hip.step(t, h, 1e-12)
        t = t + hdid
        hd.append(hdid)
    x, v = hip.download()
    xo, vo, hdo, hno, nfo = oracle.ra15_run(s, 120, h0=-4.0, tol=1e-12, pn=1)
    assert np.all(np.array(hd) < 0)
    assert np.array_equal(np.array(hd), hdo) and h == hno
    check(x, v, xo, vo, "RA15 h < 0")


@pytest.mark.parametrize("algor", [2, 3])
def test_bs_backward(hip, oracle, algor):
    s = systems.config_small(seed=13, nparticles=600)
    say(f"case: algorithm {algor} backward (first trial -4 d), 8 planets + 600 particles, 25 steps")
    hip.configure(algor)
    hip.upload_system(s)
    h, t, hd, nr = -4.0, 0.0, [], 0
    for _ in range(25):
        h, hdid, nf, r = hip.step(t, h, 1e-12)
        t = t + hdid
        hd.append(hdid)
        nr += r
    x, v = hip.download()
    xo, vo, hdo, hno, rej, _ = oracle.bs_run(algor, s, 25, h0=-4.0, tol=1e-12)
    assert np.all(np.array(hd) < 0)
    assert np.array_equal(np.array(hd), hdo) and h == hno and nr == rej
    assert np.array_equal(x[1:], xo[1:]) and np.array_equal(v[1:], vo[1:])


@pytest.mark.parametrize("mode", ["backward", "oblate"])
def test_hybrid_backward_and_oblate(hip, oracle, mode):
    s = systems.make_system(43, 600, a_range=(0.5, 40.0), emax=0.3, imax_deg=20.0,
                            h=-2.0 if mode == "backward" else 2.0, near_fraction=0.5)
    s["jcen"] = np.array(JC) if mode == "oblate" else np.zeros(3)
    say(f"case: HYBRID {mode}, 8 planets + 600 particles (half near planet orbits), 60 steps")
    x, v, rows, most = hy_run_device(hip, oracle, s, 60)
    xo, vo, evi, evd, nf = oracle.hy_run(s, 60)
    say(f"encounter pairs in one step, at most {most}; oracle events {len(evi)}")
    assert most > 0
    check(x, v, xo, vo, f"HYBRID {mode}")
    compare_rows(rows, oracle_rows(evi, evd))


def line_of_particles(count):
    """One low-mass planet with a huge encounter radius and `count` particles
    in a short line beside it: every particle is a close encounter in the
    same step (the reference's buffer-growth case, tests/test_events.py:19-30)."""
    mu = systems.K2
    m = np.concatenate([[mu, 1e-12 * mu], np.zeros(count)])
    x = np.zeros((count + 2, 3)); v = np.zeros((count + 2, 3))
    x[1] = [1.0, 0.0, 0.0]; v[1] = [0.0, np.sqrt(mu), 0.0]
    x[2:, 0] = 1.001 + np.arange(count) * 1e-10
    x[2:, 1] = -0.005
    v[2:, 1] = 0.1
    s = {"m": m, "x": x, "v": v, "nbig": 2, "h": 0.1, "rcen": 0.005, "jcen": np.zeros(3)}
    systems.finish_limits(s)
    s["rce"][1] = 0.5
    return s


def test_event_list_grows_within_a_step(hip, oracle, monkeypatch):
    s = line_of_particles(4100)
    monkeypatch.setenv("MERCURY_HIP_EVENT_CAP", "1000")
    hip.free()
    say("case: 4100 encounters in one step against an event list of 1000")
    x, v, rows = mvs_stepwise(hip, s, 1)
    hip.free()
    xo, vo, evi, evd = oracle.mvs_run(s, 1)
    say(f"events: device {len(rows)}, oracle {len(evi)}")
    assert len(evi) >= 4100
    compare_rows(rows, oracle_rows(evi, evd))
    check(x, v, xo, vo, "buffer growth")


def test_advance_event_overflow_fails_loudly(hip, monkeypatch):
    from mercuda_amd.hip import MercuryHipError
    s = line_of_particles(4100)
    monkeypatch.setenv("MERCURY_HIP_EVENT_CAP", "1000")
    hip.free()
    hip.configure(1)
    hip.upload_system(s)
    with pytest.raises(MercuryHipError):
        hip.advance(0.0, s["h"], 2, s["rcen"], screen=True)
    hip.free()


def test_radau_100k_particles(hip, oracle):
    s = systems.config_large(seed=2, nparticles=100_000)
    say("case: RA15 tol 1e-12 with 1PN, 8 planets + 100000 particles, 8 sequences")
    hip.configure(4)
    hip.upload_system(s, pn=1)
    h, t, hd = s["h"], 0.0, []
    for _ in range(8):
        h, hdid, nf, nr = hip.step(t, h, 1e-12)
        t = t + hdid
        hd.append(hdid)
    x, v = hip.download()
    xo, vo, hdo, hno, nfo = oracle.ra15_run(s, 8, tol=1e-12, pn=1)
    say(f"sizes {np.round(hd, 6).tolist()}")
    assert np.array_equal(np.array(hd), hdo) and h == hno
    check(x, v, xo, vo, "RA15 100k")


def test_ensemble_512_systems_sampled(hip, oracle):
    ens = systems.ensemble(seed=4)
    nsteps = 24
    say(f"case: ensemble of {len(ens)} systems, {sum(len(q['m']) for q in ens)} bodies, {nsteps} steps; "
        "16 systems checked")
    hip.configure(1)
    hip.ensemble_upload(ens)
    hip.set_tuning(steps_per_launch=16)
    hip.ensemble_advance(0.0, nsteps, screen=True)
    states = hip.ensemble_download()
    rng = np.random.default_rng(5)
    for k in sorted(rng.choice(len(ens), 16, replace=False)):
        q = ens[k]
        xo, vo, _, _ = oracle.mvs_run(q, nsteps)
        ex, _ = rel_err(states[k][0][1:], xo[1:])
        ev, _ = rel_err(states[k][1][1:], vo[1:])
        assert ex <= REL_TOL and ev <= REL_TOL, (k, ex, ev)


def test_hybrid_50k_particles(hip, oracle):
    from mercuda_amd import hybrid
    s = systems.config_encounters(seed=3, nparticles=50_000)
    say("case: HYBRID, 8 planets + 50000 particles (half near planet orbits), 2 steps of 2 d, device loop")
    hybrid.mdt_hkce.substeps = 0
    x, v, most = hybrid.hy_run(hip, s, 2)
    say(f"encounter pairs in one step, at most {most}; BS2 substeps {hybrid.mdt_hkce.substeps}")
    xo, vo, _, _, _ = oracle.hy_run(s, 2)
    assert most > 1000
    check(x, v, xo, vo, "HYBRID 50k")
