"""Parity of the HIP HYBRID path (algorithm 10) with the oracle's
restatement of mdt_hy (mercury6_2.for:3618-3797), whose own parity with the
reference is bitwise (tests/test_oracle.py): the regular map on the device
(kick, Sun-momentum drift, Kepler drift, mce_snif scan, mfo_hy), encounter
members integrated by the device's BS2 (mdt_bs2 with mfo_hkce), the driver's
mdt_hkce loop and event handling as in tests/hybrid_driver.py."""
import numpy as np
import pytest

from helpers import REL_TOL, compare_rows, oracle_rows, rel_err
from hybrid_driver import hy_run_device
from mercuda_amd import systems

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def hybrid_system(seed, nparticles, near=0.5, h=2.0):
    s = systems.make_system(seed, nparticles, a_range=(0.5, 40.0), emax=0.3, imax_deg=20.0, h=h,
                            near_fraction=near)
    s["jcen"] = np.zeros(3)
    return s


def check(x, v, xo, vo, label):
    ex, nx = rel_err(x[1:], xo[1:])
    ev, nv = rel_err(v[1:], vo[1:])
    say(f"{label}: max rel err x {ex:.3e} v {ev:.3e}; bodies differing {nx}/{len(x) - 1}")
    assert ex <= REL_TOL and ev <= REL_TOL


def test_bs2_compact_matches_oracle(hip, oracle):
    """One planet and 40 particles inside its changeover radius, plus a second
    planet: the device BS2 step (mdt_bs2 with mfo_hkce) against the oracle's."""
    s = hybrid_system(31, 0, near=0.0)
    rng = np.random.default_rng(3)
    m = np.concatenate([s["m"][:3], np.zeros(40)])
    x = np.zeros((43, 3)); v = np.zeros((43, 3))
    x[1:3] = s["x"][1:3]; v[1:3] = s["v"][1:3]
    rc = np.zeros(43)
    rc[1:3] = 0.05
    rc[3:] = 0.01
    x[3:] = x[1] + rng.normal(0, 0.02, (40, 3))
    v[3:] = v[1] + rng.normal(0, 2e-3, (40, 3))
    pairs = [(2, 3)] + [(2, k) for k in range(4, 44)]
    hip.configure(10)
    hip.upload_system(dict(s, m=s["m"], x=s["x"], v=s["v"]))
    for h0 in (0.5, 2.0, 40.0):
        hip.encounter_enter(m, x, v, rc, np.zeros(43), np.zeros(43), 3, pairs)
        hn, hd, nf, nr = hip.step(0.0, h0, 1e-12)
        xd, vd = hip.encounter_download()
        hip.encounter_exit()
        xo, vo, hdo, hno, rej, nfo = oracle.bs2(m, x, v, rc, pairs, h0, tol=1e-12)
        say(f"h0 {h0}: device hdid {hd} next {hn} halvings {nr}; oracle {hdo} {hno} {rej}")
        assert hd == hdo and hn == hno and nr == rej
        assert np.array_equal(xd[1:], xo[1:]) and np.array_equal(vd[1:], vo[1:])


@pytest.mark.parametrize("variant", ["fused", "split", "paired"])
@pytest.mark.parametrize("nparticles", [0, 40])
def test_device_encounter_loop_compact_sets(hip, variant, nparticles, monkeypatch):
    """The device encounter loop on hand-made sets with a Big-Big pair: two
    planets inside each other's changeover radius, with no particles (a set of
    Big members only) or with 40 particles around the first.  Same substeps,
    step recommendation and bitwise states as mdt_hkce's loop of one
    mercury_hip_step per substep (mercury6_2.for:3883-3905)."""
    import math
    monkeypatch.setenv("MERCURY_HIP_ENC_LOOP", variant)
    s = hybrid_system(31, 0, near=0.0)
    rng = np.random.default_rng(5)
    n = 3 + nparticles
    m = np.concatenate([s["m"][:3], np.zeros(nparticles)])
    x = np.zeros((n, 3)); v = np.zeros((n, 3))
    x[1] = s["x"][1]; v[1] = s["v"][1]
    x[2] = x[1] + np.array([0.01, -0.004, 0.002]); v[2] = v[1] + np.array([1e-4, 2e-4, -5e-5])
    rc = np.zeros(n)
    rc[1:3] = 0.05
    rc[3:] = 0.01
    if nparticles:
        x[3:] = x[1] + rng.normal(0, 0.02, (nparticles, 3))
        v[3:] = v[1] + rng.normal(0, 2e-3, (nparticles, 3))
    pairs = [(2, 3)] + [(2, k) for k in range(4, n + 1)]
    hip.configure(10)
    hip.upload_system(s)
    h0, hrec0, tol = 40.0, 40.0, 1e-12
    say(f"case: compact encounter set, 2 planets + {nparticles} particles, loop {variant}, h0 {h0}")
    hip.encounter_enter(m, x, v, rc, np.zeros(n), np.zeros(n), 3, pairs)
    hrec_d, subs_d, _, _ = hip.encounter_advance(h0, hrec0, tol)
    xd, vd = hip.encounter_download()
    hip.encounter_exit()
    hip.encounter_enter(m, x, v, rc, np.zeros(n), np.zeros(n), 3, pairs)
    tlocal, hlocal, hrec, subs = 0.0, math.copysign(hrec0, h0), hrec0, 0
    while True:
        tmp0 = abs(h0) - abs(tlocal)
        hrec = hlocal
        if abs(hlocal) > tmp0:
            hlocal = math.copysign(tmp0, h0)
        hlocal, hdid, _, _ = hip.step(tlocal, hlocal, tol)
        tlocal += hdid
        subs += 1
        if not ((tlocal - h0) * h0 < 0.0):
            break
    xh, vh = hip.encounter_download()
    hip.encounter_exit()
    say(f"substeps device {subs_d}, per call {subs}; recommendation {hrec_d} / {hrec}")
    assert subs_d == subs > 1 and hrec_d == hrec
    assert np.array_equal(xd, xh) and np.array_equal(vd, vh)


def test_hybrid_encounters_match_oracle(hip, oracle):
    s = hybrid_system(3, 300)
    say("case: HYBRID, 8 planets + 300 particles (half near planet orbits), 150 steps of 2 d")
    x, v, rows, most = hy_run_device(hip, oracle, s, 150)
    xo, vo, evi, evd, nf = oracle.hy_run(s, 150)
    say(f"encounter pairs in one step, at most {most}; oracle events {len(evi)}")
    assert most > 0
    check(x, v, xo, vo, "HYBRID encounters")
    compare_rows(rows, oracle_rows(evi, evd))


def test_hybrid_regular_map_matches_oracle(hip, oracle):
    s = hybrid_system(4, 3000, near=0.0)
    say("case: HYBRID regular map, 8 planets + 3000 particles, 200 steps of 2 d")
    x, v, rows, most = hy_run_device(hip, oracle, s, 200)
    xo, vo, evi, evd, nf = oracle.hy_run(s, 200)
    check(x, v, xo, vo, "HYBRID regular")
    compare_rows(rows, oracle_rows(evi, evd))


@pytest.mark.parametrize("count", [0, 1, 127, 128, 129])
def test_hybrid_ragged_counts(hip, oracle, count):
    s = hybrid_system(11, count, near=0.5)
    say(f"case: HYBRID ragged count {count}, 40 steps")
    x, v, rows, most = hy_run_device(hip, oracle, s, 40)
    xo, vo, evi, evd, nf = oracle.hy_run(s, 40)
    check(x, v, xo, vo, f"HYBRID count {count}")
    compare_rows(rows, oracle_rows(evi, evd))


@pytest.mark.parametrize("variant", ["fused", "split", "paired"])
def test_device_encounter_loop_equals_per_substep_calls(hip, oracle, variant, monkeypatch):
    """mercury_hip_encounter_advance (mdt_hkce's loop with the step control on
    the device) against the same loop driven by one mercury_hip_step per
    substep: the same substeps and bitwise the same states; and the oracle.
    Both device variants: one launch per attempt (every block keeps the Big
    members itself), and the Big members' launch followed by the particles'."""
    from mercuda_amd import hybrid
    monkeypatch.setenv("MERCURY_HIP_ENC_LOOP", variant)
    s = hybrid_system(5, 2000)
    say(f"case: HYBRID encounter loop on the device ({variant}) vs per-substep calls, 8 planets + 2000 particles, "
        "60 steps")
    hybrid.mdt_hkce.substeps = 0
    xd, vd, most = hybrid.hy_run(hip, s, 60)
    nd = hybrid.mdt_hkce.substeps
    hybrid.mdt_hkce.substeps = 0
    xh, vh, _ = hybrid.hy_run(hip, s, 60, device_loop=False)
    nh = hybrid.mdt_hkce.substeps
    say(f"encounter pairs in one step, at most {most}; BS2 substeps device {nd}, per-call {nh}")
    assert most > 0 and nd == nh > 0
    assert np.array_equal(xd, xh) and np.array_equal(vd, vh)
    xo, vo, _, _, _ = oracle.hy_run(s, 60)
    check(xd, vd, xo, vo, "HYBRID device encounter loop")
