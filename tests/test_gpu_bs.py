"""Parity of the HIP Bulirsch-Stoer paths, algorithm 2 (mdt_bs1) and 3
(mdt_bs2), each with mfo_all (mercury6_2.for:3227-3617, 4506-4707), with the
oracle's restatement, whose own parity with the reference is bitwise
(tests/test_oracle.py).  The step control runs on the host with the Fortran's
rules, so when the device levels reproduce the Fortran arithmetic the step
sizes and the states agree bit for bit."""
import numpy as np
import pytest

from helpers import compare_rows, hip_event_rows
from mercuda_amd import systems

pytestmark = pytest.mark.gpu

JC = (1e-6 * 0.0046 ** 2, -2e-7 * 0.0046 ** 4, 5e-8 * 0.0046 ** 6)


def say(*a):
    print(*a, flush=True)


def bs_steps(hip, algor, sysd, nsteps, *, tol=1e-12, pn=0, h0=None):
    """mercury_hip_step loop (the drop-in path): states, sizes done, next
    proposal, halvings."""
    hip.configure(algor)
    hip.upload_system(sysd, pn=pn)
    h = sysd["h"] if h0 is None else h0
    time = 0.0
    hd, nr = [], 0
    for _ in range(nsteps):
        h, hdid, nf, r = hip.step(time, h, tol)
        time = time + hdid
        hd.append(hdid)
        nr += r
    x, v = hip.download()
    return x, v, np.array(hd), h, nr


def check_bitwise(hip, oracle, algor, sysd, nsteps, label, **kw):
    x, v, hd, hn, nr = bs_steps(hip, algor, sysd, nsteps, **kw)
    xo, vo, hdo, hno, rej, _ = oracle.bs_run(algor, sysd, nsteps, h0=kw.get("h0"), tol=kw.get("tol", 1e-12),
                                             pn=kw.get("pn", 0))
    say(f"{label}: sizes first {hd[:3]} last {hd[-1]:.5f}; halvings {nr} vs oracle {rej}; "
        f"bodies differing x {int(np.sum(np.any(x != xo, axis=1)))} v {int(np.sum(np.any(v != vo, axis=1)))}")
    assert np.array_equal(hd, hdo) and hn == hno and nr == rej
    assert np.array_equal(x[1:], xo[1:]) and np.array_equal(v[1:], vo[1:])


@pytest.mark.parametrize("algor", [2, 3])
@pytest.mark.parametrize("pn,obl", [(0, False), (1, True)])
def test_bs_matches_oracle(hip, oracle, algor, pn, obl):
    s = systems.config_small(seed=21, nparticles=700)
    if obl:
        s["jcen"] = np.array(JC)
    say(f"case: algorithm {algor} pn={pn} obl={obl}, 8 planets + 700 particles, 40 steps")
    check_bitwise(hip, oracle, algor, s, 40, f"BS{algor}", pn=pn, h0=6.0)


@pytest.mark.parametrize("algor", [2, 3])
def test_bs_halvings(hip, oracle, algor):
    """A first step far too large: the attempts halve it, as the Fortran's."""
    s = systems.config_small(seed=22, nparticles=300)
    say(f"case: algorithm {algor}, first step 400 d, 15 steps")
    check_bitwise(hip, oracle, algor, s, 15, f"BS{algor} halvings", h0=400.0, pn=1)


@pytest.mark.parametrize("count", [0, 1, 15, 16, 17, 129])
def test_bs_ragged_counts(hip, oracle, count):
    s = systems.config_small(seed=23, nparticles=count)
    say(f"case: algorithms 2 and 3, ragged count {count}, 12 steps")
    for algor in (2, 3):
        check_bitwise(hip, oracle, algor, s, 12, f"BS{algor} count {count}", h0=5.0)


def wide_system(npl, nparticles, seed):
    """npl low-mass planets on near-circular orbits spaced by a factor 1.3 in a
    (0.6 AU outward), with particles among them: more planets than the Hill-
    stable draw of systems.make_system places."""
    rng = np.random.default_rng(seed)
    base = systems.make_system(seed, nparticles, nplanets=0, h=4.0, a_range=(0.5, 30.0))
    a = 0.6 * 1.3 ** np.arange(npl)
    pm = 10 ** rng.uniform(-8, -6, npl) * systems.K2
    e = rng.uniform(0, 0.02, npl); inc = rng.uniform(0, 1, npl) * systems.DEG
    w, node, M = rng.uniform(0, 2 * np.pi, (3, npl))
    xp, vp = systems.elements_to_xv(base["m"][0] + pm, a, e, inc, w, node, M)
    m = np.concatenate([base["m"][:1], pm, base["m"][1:]])
    x = np.concatenate([base["x"][:1], xp, base["x"][1:]])
    v = np.concatenate([base["v"][:1], vp, base["v"][1:]])
    s = dict(m=m, x=x, v=v, nbig=1 + npl, h=4.0, rcen=base["rcen"], jcen=np.zeros(3), seed=seed)
    return systems.finish_limits(s)


def test_bs_planet_counts(hip, oracle):
    """2, 12 and 16 planets (one level's lanes), and no planets at all."""
    for npl in (0, 2, 12, 16):
        s = wide_system(npl, 200, 24 + npl)
        say(f"case: {npl} planets + 200 particles, both algorithms, 10 steps")
        for algor in (2, 3):
            check_bitwise(hip, oracle, algor, s, 10, f"BS{algor} {npl} planets")


@pytest.mark.parametrize("algor", [2, 3])
@pytest.mark.parametrize("npl", [17, 40, 63])
def test_bs_more_than_16_planets(hip, oracle, algor, npl):
    """More planets than one 16-lane level holds, up to the 63 a system may
    carry: the levels run 64 lanes wide, bitwise the oracle's states and sizes."""
    s = wide_system(npl, 200, 40 + npl)
    say(f"case: algorithm {algor}, {npl} planets + 200 particles, 1PN, 12 steps")
    check_bitwise(hip, oracle, algor, s, 12, f"BS{algor} {npl} planets", pn=1, h0=2.0)


def test_bs_too_many_big_bodies_fails_loudly(hip):
    """64 planets besides the central body are more than the library holds:
    the upload refuses them."""
    s = wide_system(64, 50, 40)
    hip.configure(3)
    with pytest.raises(Exception):
        hip.upload_system(s)


def test_bs_events_match_mce_stat(hip, oracle):
    """events() after each BS step = mce_stat + mce_cent (mal_hvar,
    mercury6_2.for:452-515) on the same step-start and end states."""
    s = systems.config_encounters(seed=5, nparticles=400)
    s["rce"][1:s["nbig"]] *= 3.0
    rcen = 0.05
    nb = s["nbig"]
    total = 0
    for algor in (2, 3):
        say(f"case: algorithm {algor} events, 8 planets + 400 particles near orbits, 30 steps")
        hip.configure(algor)
        hip.upload_system(s)
        h, time = 2.0, 0.0
        for step in range(30):
            h, hdid, _, _ = hip.step(time, h, 1e-12)
            time = time + hdid
            ev = hip.events(time, hdid, rcen)
            x0, v0 = hip.download(previous=True)
            x1, v1 = hip.download()
            got = hip_event_rows(ev, step)
            el = oracle.mce_stat(time, hdid, s["m"], x0, v0, x1, v1, s["rce"], s["rphys"], nb)
            want = [(step, 1, i, j, d, t) for (i, j, d, t) in el.clo()]
            want += [(step, 6 if c == 1 else 2, i, j, d, t) for (i, j, c, d, t) in el.hit()]
            want += [(step, 8, 1, j, d, t) for (j, t, d) in oracle.mce_cent(time, hdid, rcen, s["m"], x0, v0, x1, v1)]
            compare_rows(got, want, rtol=1e-12)
            total += len(want)
    say(f"events compared: {total}")
    assert total > 10

