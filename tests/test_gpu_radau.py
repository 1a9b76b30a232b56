"""Parity of the HIP RA15 path (algorithm 4) with the oracle's restatement of
mdt_ra15 + mfo_all (mercury6_2.for:4133-4707) on identical seeded inputs.

The step-size law runs on the host with the same libm pow as the CPU path,
so when the device iteration reproduces the Fortran arithmetic the sequence
of step sizes is identical and states agree to rounding.  The reference's own
RADAU checks are the CPU-vs-GPU comparison of tests/test_gpu_algorithms.py
and the 1e-12 circular-orbit answer of tests/test_release.py:35-48."""
import numpy as np
import pytest

from helpers import REL_TOL, compare_rows, hip_event_rows, rel_err
from mercuda_amd import systems

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def radau_steps(hip, sysd, nsteps, *, tol=1e-12, pn=0, h0=None, reset_at=(), screen_rcen=None):
    """mercury_hip_step loop (the drop-in path): states, sizes done, next
    proposal, force count, redo count and (optionally) per-step event rows
    together with the step-start/end states they were screened on."""
    hip.configure(4)
    hip.upload_system(sysd, pn=pn)
    h = sysd["h"] if h0 is None else h0
    time = 0.0
    hdid_all, nf_all, nr_all, steps = [], 0, 0, []
    for s in range(nsteps):
        if s in reset_at:
            hip.reset(0)
        h, hdid, nf, nr = hip.step(time, h, tol)
        time = time + hdid
        hdid_all.append(hdid)
        nf_all += nf
        nr_all += nr
        if screen_rcen is not None:
            ev = hip.events(time, hdid, screen_rcen)
            x0, v0 = hip.download(previous=True)
            x1, v1 = hip.download()
            steps.append((s, time, hdid, ev, x0, v0, x1, v1))
    x, v = hip.download()
    return x, v, np.array(hdid_all), h, nf_all, nr_all, steps


def check_state(x, v, xo, vo, label, tol=REL_TOL):
    ex, nx = rel_err(x[1:], xo[1:])
    ev, nv = rel_err(v[1:], vo[1:])
    say(f"{label}: max rel err x {ex:.3e} v {ev:.3e}; bodies differing {nx}/{len(x) - 1}")
    assert ex <= tol and ev <= tol


@pytest.mark.parametrize("pn", [0, 1])
@pytest.mark.parametrize("obl", [False, True])
def test_force_matches_mfo_all(hip, oracle, pn, obl):
    """mercury_hip_force = mfo_all (mfo_grav with 1PN, plus mfo_obl) of the
    uploaded state, body 0 zero."""
    sysd = systems.config_small(seed=4, nparticles=3000)
    if obl:
        sysd["jcen"] = (1e-6 * 0.0046 ** 2, -2e-7 * 0.0046 ** 4, 5e-8 * 0.0046 ** 6)
    say(f"case: force pn={pn} obl={obl}, 8 planets + 3000 particles")
    hip.configure(4)
    hip.upload_system(sysd, pn=pn)
    a = hip.force()
    ao = oracle.mfo_all(sysd, pn=pn)
    err, ndiff = rel_err(a[1:], ao[1:])
    say(f"max rel err {err:.3e}, bodies differing {ndiff}")
    assert np.all(a[0] == 0.0)
    assert err <= 1e-15


@pytest.mark.parametrize("pn", [0, 1])
def test_radau_matches_oracle(hip, oracle, pn):
    sysd = systems.config_small(seed=1, nparticles=1000)
    say(f"case: RA15 pn={pn}, 8 planets + 1000 particles, 300 sequences, tol 1e-12")
    x, v, hd, hn, nf, nr, _ = radau_steps(hip, sysd, 300, pn=pn)
    xo, vo, hdo, hno, nfo = oracle.ra15_run(sysd, 300, tol=1e-12, pn=pn)
    say(f"sizes: first {hd[:3]}, last {hd[-1]:.6f}; force evaluations {nf} vs oracle {nfo}; redo {nr}")
    assert np.array_equal(hd, hdo) and hn == hno
    assert nf == nfo
    check_state(x, v, xo, vo, f"RA15 pn={pn}")


def test_radau_redo_on_reset(hip, oracle):
    """A trial size far too large: the reset sequence shrinks by 0.8 and
    redoes (mercury6_2.for:4413-4418) until it stops shrinking."""
    sysd = systems.config_small(seed=3, nparticles=500)
    say("case: RA15 first trial 200 d, 40 sequences")
    x, v, hd, hn, nf, nr, _ = radau_steps(hip, sysd, 40, h0=200.0, pn=1)
    xo, vo, hdo, hno, nfo = oracle.ra15_run(sysd, 40, h0=200.0, tol=1e-12, pn=1)
    say(f"redo passes {nr}; first size done {hd[0]:.6f}")
    assert nr > 0
    assert np.array_equal(hd, hdo) and nf == nfo
    check_state(x, v, xo, vo, "RA15 redo")


def test_radau_reset_mid_run(hip, oracle):
    """reset(flag != 2) discards the b/e history: the same as a fresh
    mdt_ra15 sequence (dtflag = 0) from the state reached."""
    sysd = systems.config_small(seed=5, nparticles=400)
    say("case: RA15 reset after 60 of 120 sequences")
    x, v, hd, hn, nf, nr, _ = radau_steps(hip, sysd, 120, reset_at=(60,), pn=1)
    x1, v1, hd1, hn1, _ = oracle.ra15_run(sysd, 60, tol=1e-12, pn=1)
    mid = dict(sysd); mid["x"] = x1; mid["v"] = v1
    x2, v2, hd2, hn2, _ = oracle.ra15_run(mid, 60, h0=hn1, tol=1e-12, pn=1, time0=float(np.sum(hd1)))
    assert np.array_equal(hd, np.concatenate([hd1, hd2]))
    check_state(x, v, x2, v2, "RA15 reset")


@pytest.mark.parametrize("count", [0, 1, 127, 128, 129, 4097])
def test_radau_ragged_counts(hip, oracle, count):
    sysd = systems.config_small(seed=11, nparticles=count)
    say(f"case: RA15 ragged count {count}, 30 sequences")
    x, v, hd, hn, nf, nr, _ = radau_steps(hip, sysd, 30, pn=1)
    xo, vo, hdo, hno, nfo = oracle.ra15_run(sysd, 30, tol=1e-12, pn=1)
    assert np.array_equal(hd, hdo)
    check_state(x, v, xo, vo, f"RA15 count {count}")


def test_radau_circular_orbit_known_answer(hip):
    """Test particles on exact circular orbits of the central body alone: after
    1000 d the position equals r(cos wt, sin wt) to 1e-12 AU (the reference's
    known answer, tests/test_release.py:35-48)."""
    mu = systems.K2
    r = np.array([0.7, 1.0, 1.6, 3.3])
    th0 = np.array([0.0, 1.0, 2.5, 4.0])
    k = len(r)
    m = np.concatenate([[mu], np.zeros(k)])
    x = np.zeros((k + 1, 3)); v = np.zeros((k + 1, 3))
    w = np.sqrt(mu / r ** 3)
    x[1:, 0] = r * np.cos(th0); x[1:, 1] = r * np.sin(th0)
    v[1:, 0] = -r * w * np.sin(th0); v[1:, 1] = r * w * np.cos(th0)
    sysd = {"m": m, "x": x, "v": v, "nbig": 1, "h": 10.0, "rce": np.zeros(k + 1), "rphys": np.zeros(k + 1)}
    hip.configure(4)
    hip.upload_system(sysd)
    time, h = 0.0, 10.0
    while time < 1000.0:
        h = min(h, 1000.0 - time)
        hnext, hdid, _, _ = hip.step(time, h, 1e-12)
        time = time + hdid
        if hnext != h:
            h = hnext
    xe, _ = hip.download()
    want = np.stack([r * np.cos(th0 + w * time), r * np.sin(th0 + w * time)], 1)
    err = np.abs(xe[1:, :2] - want).max()
    say(f"circular orbits after {time} d: max |dx| {err:.3e} AU")
    assert abs(time - 1000.0) < 1e-9 and err < 1e-12


def test_radau_events_match_mce_stat(hip, oracle):
    """events() after each RA15 sequence = mce_stat + mce_cent (mal_hvar,
    mercury6_2.for:452-515) on the same step-start and end states."""
    sysd = systems.config_encounters(seed=3, nparticles=1500)
    sysd["rce"][1:sysd["nbig"]] *= 3.0
    rcen = 0.05
    say("case: RA15 encounter screening, 1500 particles, 80 sequences")
    *_, steps = radau_steps(hip, sysd, 80, pn=1, h0=2.0, screen_rcen=rcen)
    nb = sysd["nbig"]
    total = 0
    for (s, time, hdid, ev, x0, v0, x1, v1) in steps:
        got = hip_event_rows(ev, s)
        el = oracle.mce_stat(time, hdid, sysd["m"], x0, v0, x1, v1, sysd["rce"], sysd["rphys"], nb)
        want = [(s, 1, i, j, d, t) for (i, j, d, t) in el.clo()]
        want += [(s, 6 if c == 1 else 2, i, j, d, t) for (i, j, c, d, t) in el.hit()]  # chit: 1 hit, -1 near
        want += [(s, 8, 1, j, d, t) for (j, t, d) in oracle.mce_cent(time, hdid, rcen, sysd["m"], x0, v0, x1, v1)]
        compare_rows(got, want, rtol=1e-12)
        total += len(want)
    say(f"events compared: {total}")
    assert total > 20
