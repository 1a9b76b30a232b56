"""Parity of the device's non-gravitational forces with the oracle's
restatement of mfo_ngf and mfo_pr (mercury6_2.for:5106-5152, 5307-5357),
itself pinned bit for bit to the reference routines (tests/test_oracle.py):
in the RA15 and Bulirsch-Stoer forces (mfo_all, every body), and the cometary
terms of MVS on particles (mdt_mvs's angf, mercury6_2.for:4054, 4090; the
corrector mco_mvs2h, 2171, 2200).

Radiation pressure, Poynting-Robertson drag and the Yarkovsky term use only
+ - * / and sqrt, so those runs agree bit for bit.  The cometary terms scale by
Marsden's g(r), which needs x^y: the device computes the correctly rounded
value (pow_cr), the CPU path libm's pow, which is within 0.52 ulp and misses
the correctly rounded value for about one argument in a thousand.  The force
test therefore asks for bitwise equality on every body whose three libm pow
calls are correctly rounded (decided with 200-bit mpmath) and 1e-15 relative
on the others; whole runs with cometary terms meet the 1e-11 bar and report
how many bodies are bit-identical."""
import math

import numpy as np
import pytest

from helpers import REL_TOL, rel_err
from mercuda_amd import systems

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def libm_pow_is_cr(x, y):
    import mpmath
    from mpmath.libmp import to_float
    with mpmath.workprec(200):
        cr = to_float(mpmath.power(mpmath.mpf(x), mpmath.mpf(y))._mpf_, rnd="n")
    return math.pow(x, y) == cr


def pow_sensitive(x, ngf):
    """Bodies whose g(r) goes through a libm pow call that is not correctly
    rounded (the cometary block runs: some A nonzero, inside the cutoff or a
    large |A|)."""
    out = set()
    for j in range(1, len(x)):
        a = ngf[j, :3]
        r2 = float(x[j] @ x[j])
        if not np.any(a != 0) or not (r2 < 88.0 or np.any(np.abs(a) > 1e-7)):
            continue
        q = math.sqrt(r2) * .3561253561253561
        p2 = math.pow(q, 5.093)
        if not (libm_pow_is_cr(q, -2.15) and libm_pow_is_cr(q, 5.093) and libm_pow_is_cr(1.0 + p2, -4.6142)):
            out.add(j)
    return out


def comet_system(seed, nparticles, kinds=("a1", "a2", "a3", "yar", "beta"), big=True, fraction=0.5):
    s = systems.config_small(seed=seed, nparticles=nparticles)
    return systems.add_nongrav(s, seed=seed + 100, fraction=fraction, kinds=kinds, big=big)


@pytest.mark.parametrize("pn", [0, 1])
@pytest.mark.parametrize("kinds", [("a1", "a2", "a3", "yar", "beta"), ("yar", "beta"), ("a1", "a2", "a3")])
def test_force_matches_mfo_all(hip, oracle, pn, kinds):
    """mercury_hip_force with the model of every parameter kind, on Big
    bodies and particles; a tenth of the comets carry A values of 1e-3 so that
    the cometary term dominates and g's last bits show in the sum."""
    s = comet_system(4, 3000, kinds=kinds)
    if "a1" in kinds:
        rng = np.random.default_rng(1)
        sel = np.flatnonzero(np.any(s["ngf"][:, :3] != 0, axis=1))
        big = rng.choice(sel, size=max(1, len(sel) // 10), replace=False)
        s["ngf"][big, :3] = np.where(s["ngf"][big, :3] != 0, np.sign(s["ngf"][big, :3]) * 1e-3, 0.0)
    say(f"case: force ngflag {s['ngflag']} pn {pn} kinds {kinds}, 8 planets + 3000 particles")
    hip.configure(4)
    hip.upload_system(s, pn=pn)
    a = hip.force()
    ao = oracle.mfo_all(s, pn=pn)
    sens = pow_sensitive(s["x"], s["ngf"])
    differ = [j for j in range(1, len(a)) if not np.array_equal(a[j], ao[j])]
    say(f"bodies differing {len(differ)}; libm pow not correctly rounded for {len(sens)} bodies")
    assert set(differ) <= sens
    err, _ = rel_err(a[1:], ao[1:])
    assert err <= 1e-15
    # the model is on: the accelerations differ from pure gravity
    ag = oracle.mfo_all(dict(s, ngflag=0), pn=pn)
    assert np.sum(np.any(ag != ao, axis=1)) > 100


def radau_run(hip, s, nsteps, pn, tol=1e-12):
    hip.configure(4)
    hip.upload_system(s, pn=pn)
    h, time, hd = s["h"], 0.0, []
    for _ in range(nsteps):
        h, hdid, _, _ = hip.step(time, h, tol)
        time = time + hdid
        hd.append(hdid)
    x, v = hip.download()
    return x, v, np.array(hd), h


@pytest.mark.parametrize("pn", [0, 1])
def test_radau_pr_yarkovsky_bitwise(hip, oracle, pn):
    """RA15 with radiation pressure, PR drag and Yarkovsky (ngflag 3 with no
    cometary A): the sizes done and the states bit for bit."""
    s = comet_system(7, 800, kinds=("yar", "beta"))
    assert s["ngflag"] == 3
    say(f"case: RA15 PR + Yarkovsky pn {pn}, 8 planets + 800 particles, 120 sequences")
    x, v, hd, hn = radau_run(hip, s, 120, pn)
    xo, vo, hdo, hno, _ = oracle.ra15_run(s, 120, tol=1e-12, pn=pn)
    say(f"sizes last {hd[-1]:.6f}; bodies differing {int(np.sum(np.any(x != xo, axis=1)))}")
    assert np.array_equal(hd, hdo) and hn == hno
    assert np.array_equal(x[1:], xo[1:]) and np.array_equal(v[1:], vo[1:])


def test_radau_all_terms_match_oracle(hip, oracle):
    """RA15 with 1PN and every term (cometary A1-A3 on planets and particles,
    Yarkovsky, radiation): the 1e-11 bar."""
    s = comet_system(8, 800)
    say("case: RA15 1PN + ngflag 3 (all terms), 8 planets + 800 particles, 120 sequences")
    x, v, hd, hn = radau_run(hip, s, 120, 1)
    xo, vo, hdo, hno, _ = oracle.ra15_run(s, 120, tol=1e-12, pn=1)
    same = int(np.sum(np.all(x == xo, axis=1) & np.all(v == vo, axis=1)))
    say(f"sizes identical {np.array_equal(hd, hdo)}; bodies bit-identical {same}/{len(x)}")
    assert np.allclose(hd, hdo, rtol=1e-12, atol=0)
    ex, _ = rel_err(x[1:], xo[1:]); ev, _ = rel_err(v[1:], vo[1:])
    say(f"max rel err x {ex:.3e} v {ev:.3e}")
    assert ex <= REL_TOL and ev <= REL_TOL


def bs_run(hip, algor, s, nsteps, pn=0, tol=1e-12):
    hip.configure(algor)
    hip.upload_system(s, pn=pn)
    h, time, hd, nr = s["h"], 0.0, [], 0
    for _ in range(nsteps):
        h, hdid, _, r = hip.step(time, h, tol)
        time = time + hdid
        hd.append(hdid)
        nr += r
    x, v = hip.download()
    return x, v, np.array(hd), h, nr


@pytest.mark.parametrize("algor", [2, 3])
def test_bs_pr_yarkovsky_bitwise(hip, oracle, algor):
    s = comet_system(9, 500, kinds=("yar", "beta"))
    say(f"case: algorithm {algor} PR + Yarkovsky pn 1, 8 planets + 500 particles, 30 steps")
    x, v, hd, hn, nr = bs_run(hip, algor, s, 30, pn=1)
    xo, vo, hdo, hno, rej, _ = oracle.bs_run(algor, s, 30, tol=1e-12, pn=1)
    assert np.array_equal(hd, hdo) and hn == hno and nr == rej
    assert np.array_equal(x[1:], xo[1:]) and np.array_equal(v[1:], vo[1:])


@pytest.mark.parametrize("algor,kinds", [(2, ("a1", "a2", "a3", "yar", "beta")), (3, ("a1", "a2", "a3"))])
def test_bs_cometary_match_oracle(hip, oracle, algor, kinds):
    s = comet_system(10, 500, kinds=kinds)
    say(f"case: algorithm {algor} ngflag {s['ngflag']}, 8 planets + 500 particles, 30 steps")
    x, v, hd, hn, nr = bs_run(hip, algor, s, 30)
    xo, vo, hdo, hno, rej, _ = oracle.bs_run(algor, s, 30, tol=1e-12)
    same = int(np.sum(np.all(x == xo, axis=1) & np.all(v == vo, axis=1)))
    say(f"sizes identical {np.array_equal(hd, hdo)}; bodies bit-identical {same}/{len(x)}")
    assert np.allclose(hd, hdo, rtol=1e-12, atol=0)
    ex, _ = rel_err(x[1:], xo[1:]); ev, _ = rel_err(v[1:], vo[1:])
    assert ex <= REL_TOL and ev <= REL_TOL


@pytest.mark.parametrize("mode", ["advance", "stepwise"])
def test_mvs_comets_match_oracle(hip, oracle, mode):
    """MVS with cometary particles (A1-A3 on half the particles): they run as
    permanent slow-set members carrying angf; the rest in the fast kernel.
    Both the fused advance (in several calls) and one step per call; then the
    corrector (export physical) against mco_mvs2h with angf."""
    s = comet_system(11, 3000, kinds=("a1", "a2", "a3"), big=False)
    ncom = int(np.sum(np.any(s["ngf"] != 0, axis=1)))
    nsteps = 400 if mode == "advance" else 120
    say(f"case: MVS {mode}, 8 planets + 3000 particles ({ncom} comets), {nsteps} steps of 8 d")
    hip.configure(1)
    hip.upload_system(s)
    if mode == "advance":
        hip.set_tuning(steps_per_launch=16)
        for k in range(4):
            hip.advance(k * 100 * s["h"], s["h"], 100, s["rcen"], screen=True)
    else:
        time = 0.0
        for _ in range(nsteps):
            hip.step(time, s["h"])
            time = time + s["h"]
    x, v = hip.download()
    xo, vo, _, _ = oracle.mvs_run(s, nsteps)
    same = int(np.sum(np.all(x == xo, axis=1) & np.all(v == vo, axis=1)))
    ex, _ = rel_err(x[1:], xo[1:]); ev, _ = rel_err(v[1:], vo[1:])
    say(f"max rel err x {ex:.3e} v {ev:.3e}; bodies bit-identical {same}/{len(x)}")
    assert ex <= REL_TOL and ev <= REL_TOL
    xh, vh = hip.export(s["h"], physical=True)
    xho, vho = oracle.mvs2h(s, xo, vo)
    eh, _ = rel_err(xh[1:], xho[1:]); evh, _ = rel_err(vh[1:], vho[1:])
    say(f"corrector: max rel err x {eh:.3e} v {evh:.3e}")
    assert eh <= REL_TOL and evh <= REL_TOL
    # without the model the comets end elsewhere
    xg, _, _, _ = oracle.mvs_run(dict(s, ngflag=0), nsteps)
    assert rel_err(xg[1:], xo[1:])[0] > 1e-9


def test_mvs_comets_reset_and_reupload(hip, oracle):
    """A reset (dtflag 0) recomputes angf from the state; a second upload
    without the model returns every particle to the fast kernel."""
    s = comet_system(12, 1000, kinds=("a1", "a2", "a3"), big=False)
    say("case: MVS comets, 60 steps, reset, 60 steps; then the same system without the model")
    hip.configure(1)
    hip.upload_system(s)
    hip.advance(0.0, s["h"], 60, s["rcen"], screen=False)
    hip.reset(0)
    hip.advance(60 * s["h"], s["h"], 60, s["rcen"], screen=False)
    x, v = hip.download()
    x1, v1, _, _ = oracle.mvs_run(s, 60, screen=False)
    mid = dict(s, x=x1, v=v1)
    xo, vo, _, _ = oracle.mvs_run(mid, 60, time0=60 * s["h"], screen=False)
    ex, _ = rel_err(x[1:], xo[1:])
    assert ex <= REL_TOL
    plain = dict(s, ngflag=0)
    plain.pop("ngf")
    hip.upload_system(plain)
    hip.advance(0.0, s["h"], 60, s["rcen"], screen=False)
    x, _ = hip.download()
    xo, _, _, _ = oracle.mvs_run(plain, 60, screen=False)
    assert rel_err(x[1:], xo[1:])[0] <= REL_TOL


@pytest.mark.parametrize("algor", [4, 1])
def test_undefined_direction_fails(hip, algor, capfd):
    """A comet with A2 and a purely radial velocity has no transverse
    direction: mfo_ngf stops the run ('Undefined comet transverse
    direction'); the device step fails with the reason."""
    from mercuda_amd.hip import MercuryHipError
    s = comet_system(13, 50, kinds=("a1",), big=False, fraction=0.0)
    j = s["nbig"] + 3
    s["x"][j] = [2.0, 0.0, 0.0]          # (x.v)/r^2 x == v exactly: zero transverse part
    s["v"][j] = [0.003, 0.0, 0.0]
    s["ngf"][j] = [0.0, 1e-8, 0.0, 0.0, 0.0]
    s["ngflag"] = 1
    say(f"case: algorithm {algor}, comet {j} with A2 on a radial velocity")
    hip.configure(algor)
    hip.upload_system(s)
    with pytest.raises(MercuryHipError):
        if algor == 4:
            hip.step(0.0, s["h"], 1e-12)
        else:
            hip.advance(0.0, s["h"], 4, s["rcen"], screen=False)
    assert "non-gravitational force failed" in capfd.readouterr().err


@pytest.mark.parametrize("loop", ["per-substep", "device"])
def test_hybrid_comets_match_oracle(hip, oracle, loop):
    """HYBRID with the cometary terms on Big bodies and particles, half the
    particles near planet orbits (encounter members through mdt_hkce): angf at
    a reset and after the second Sun-momentum drift, in both kicks.  The
    driver's loop of one BS2 step per call, and the device encounter loop."""
    from helpers import compare_rows, oracle_rows
    from hybrid_driver import hy_run_device
    from mercuda_amd import hybrid
    s = systems.make_system(3, 300, a_range=(0.5, 40.0), emax=0.3, imax_deg=20.0, h=2.0, near_fraction=0.5)
    systems.add_nongrav(s, seed=51, fraction=0.6, kinds=("a1", "a2", "a3"), big=True)
    say(f"case: HYBRID cometary ({loop} encounter loop), 8 planets + 300 particles, 150 steps of 2 d")
    if loop == "per-substep":
        x, v, rows, most = hy_run_device(hip, oracle, s, 150)
    else:
        x, v, most = hybrid.hy_run(hip, s, 150)
        rows = None
    xo, vo, evi, evd, _ = oracle.hy_run(s, 150)
    same = int(np.sum(np.all(x == xo, axis=1) & np.all(v == vo, axis=1)))
    ex, _ = rel_err(x[1:], xo[1:]); ev, _ = rel_err(v[1:], vo[1:])
    say(f"encounter pairs in one step, at most {most}; max rel err x {ex:.3e} v {ev:.3e}; "
        f"bit-identical {same}/{len(x)}")
    assert most > 0
    assert ex <= REL_TOL and ev <= REL_TOL
    if rows is not None:
        compare_rows(rows, oracle_rows(evi, evd))
    x0, _, _, _, _ = oracle.hy_run(dict(s, ngflag=0), 150)
    assert rel_err(x0[1:], xo[1:])[0] > 1e-9
