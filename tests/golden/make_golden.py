"""Write the golden vectors under tests/golden/ from the reference's own CPU
routines (oracle/_ref/libmercury_ref.so, compiled from mercury6_2.for by
oracle/Makefile).  Run here, where /root/reference exists:

    make -C oracle ref && python tests/golden/make_golden.py [radau|hybrid]

(`radau` rewrites only the mfo_all / mdt_ra15 / mco_mvs2h fixtures, `hybrid`
only the mdt_hy runs, `bs` only the BS runs, `mce_stat` only mce_stat's
encounter and near-miss/collision lists, `nongrav` only the non-gravitational
forces and runs.)

Each fixture holds inputs and the reference's outputs only (data, no code).
"""
from __future__ import annotations

import sys
import threading
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from mercuda_amd import systems  # noqa: E402
from oracle.bindings import Reference  # noqa: E402

MU = systems.K2


def drift_cases(rng, k):
    """Random states spanning every drift_dan branch: kepmd, kepu Newton and
    Laguerre, elliptic and hyperbolic guesses, p3solve, retry path."""
    r = 10 ** rng.uniform(-1.5, 1.5, k)
    vc = np.sqrt(MU / r)
    f = rng.choice([0.1, 0.5, 0.9, 1.0, 1.2, 1.41, 1.4142135623730951, 1.5, 2.0, 3.0], k) \
        * rng.uniform(0.95, 1.05, k)
    xd = rng.normal(size=(k, 3)); xd /= np.linalg.norm(xd, axis=1)[:, None]
    vd = rng.normal(size=(k, 3)); vd /= np.linalg.norm(vd, axis=1)[:, None]
    xv = np.concatenate([xd * r[:, None], vd * (f * vc)[:, None]], axis=1)
    dt = 10 ** rng.uniform(-1, 3, k) * rng.choice([-1.0, 1.0], k)
    mu = np.full(k, MU) * rng.choice([1.0, 1.0 + 1e-3, 0.5], k)
    return mu, xv, dt


def main():
    R = Reference()
    rng = np.random.default_rng(20261203)

    # drift_one
    mu, xv, dt = drift_cases(rng, 1500)
    out = np.zeros_like(xv); flag = np.zeros(len(dt), np.int32)
    for q in range(len(dt)):
        out[q], flag[q] = R.drift_one(mu[q], xv[q], dt[q])
    np.savez_compressed(HERE / "drift_one.npz", mu=mu, xv=xv, dt=dt, out=out, flag=flag)

    # stumpff and mco_sine
    xs = rng.uniform(-80, 80, 400) * 10 ** rng.uniform(-4, 0, 400)
    st = np.array([R.stumpff(x) for x in xs])
    ang = rng.uniform(-40, 40, 400)
    sc = np.array([R.mco_sine(a) for a in ang])
    np.savez_compressed(HERE / "stumpff_sine.npz", x=xs, c=st, ang=ang, sc=sc)

    # mfo_mvs on random systems (Big + Small)
    cases = {}
    for c, (npl, npart, jc) in enumerate([(8, 40, (0, 0, 0)), (3, 10, (0, 0, 0)), (1, 5, (0, 0, 0)),
                                          (12, 30, (0, 0, 0)), (5, 20, (1e-6, -2e-9, 3e-12))]):
        s = systems.make_system(100 + c, npart, nplanets=npl)
        a = R.mfo_mvs(s["m"], s["x"], s["nbig"], jcen=jc)
        cases[f"m{c}"] = s["m"]; cases[f"x{c}"] = s["x"]; cases[f"nbig{c}"] = s["nbig"]
        cases[f"jcen{c}"] = np.array(jc, float); cases[f"a{c}"] = a
    np.savez_compressed(HERE / "mfo_mvs.npz", **cases)

    # mce_min
    k = 600
    d0 = 10 ** rng.uniform(-6, 1, k); d1 = 10 ** rng.uniform(-6, 1, k)
    d0t = rng.normal(size=k) * 10 ** rng.uniform(-6, 0, k)
    d1t = rng.normal(size=k) * 10 ** rng.uniform(-6, 0, k)
    h = rng.choice([-8.0, 2.0, 4.0, 8.0], k)
    mm = np.array([R.mce_min(d0[q], d1[q], d0t[q], d1t[q], h[q]) for q in range(k)])
    np.savez_compressed(HERE / "mce_min.npz", d0=d0, d1=d1, d0t=d0t, d1t=d1t, h=h, out=mm)

    # whole MVS runs: mdt_mvs + the driver's mce_cent, encounter-rich and plain
    runs = {}
    for c, (seed, npart, nsteps, near) in enumerate([(1, 200, 456, 0.0), (3, 300, 300, 0.5)]):
        if near:
            s = systems.make_system(seed, npart, a_range=(0.5, 40.0), emax=0.3, imax_deg=20.0, h=2.0,
                                    near_fraction=near)
            s["rce"][1:s["nbig"]] *= 3.0
        else:
            s = systems.config_small(seed, npart)
        x, v, evi, evd = R.mdt_mvs_run(s, nsteps)
        for key in ("m", "x", "v", "rce", "rphys"):
            runs[f"{key}{c}"] = s[key]
        runs[f"meta{c}"] = np.array([s["nbig"], nsteps, s["h"], s["rcen"]], float)
        runs[f"xout{c}"] = x; runs[f"vout{c}"] = v; runs[f"evi{c}"] = evi; runs[f"evd{c}"] = evd
    np.savez_compressed(HERE / "mvs_runs.npz", **runs)

    # central impacts: plunging particles through rcen
    s = systems.config_small(7, 0)
    k = 200
    r0 = rng.uniform(0.05, 0.5, k)
    q = rng.uniform(0.0, 0.012, k)
    x0 = np.zeros((k, 3)); v0 = np.zeros((k, 3))
    x0[:, 0] = r0
    # velocity for pericentre q from r0: vt = sqrt(2 mu q / (r0 (r0+q))) * sqrt(r0/q)... use energy/AM
    vt = np.sqrt(2 * MU * q / (r0 * (r0 + q)))
    vr = -np.sqrt(np.maximum(2 * MU / r0 - vt ** 2 - 2 * MU / (r0 + q) * 0, 0) * 0.9)
    v0[:, 0] = vr; v0[:, 1] = vt
    m = np.concatenate([s["m"], np.zeros(k)])
    X0 = np.concatenate([s["x"], x0]); V0 = np.concatenate([s["v"], v0])
    X1 = X0.copy(); V1 = V0.copy()
    for j in range(s["nbig"], len(m)):
        out1, _ = R.drift_one(MU, np.concatenate([X0[j], V0[j]]), 4.0)
        X1[j], V1[j] = out1[:3], out1[3:]
    hits = R.mce_cent(4.0, 4.0, 0.005, m, X0, V0, X1, V1)
    np.savez_compressed(HERE / "mce_cent.npz", m=m, x0=X0, v0=V0, x1=X1, v1=V1,
                        hits=np.array(hits, float).reshape(-1, 3))
    radau(R)
    bs(R)
    hybrid(R)
    mce_stat(R)
    nongrav(R)
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


def stat_case(rng, R, nplanets, nparticles, h, near):
    """Start and end states of one step for mce_stat: planets on near-circular
    orbits drifted by h; particles (and, in pairs, the planets after the first)
    placed near a planet with a straight relative path whose closest approach,
    as a fraction of the pair's summed physical radii, is drawn around the
    near-miss (2x) and collision (1x) thresholds."""
    nb = nplanets + 1
    n = nb + nparticles
    m = np.zeros(n)
    m[0] = MU
    m[1:nb] = 10 ** rng.uniform(-9, -5, nplanets) * MU
    rphys = np.zeros(n)
    rphys[1:] = 10 ** rng.uniform(-4.5, -3, n - 1)
    rce = np.zeros(n)
    rce[1:] = 10 ** rng.uniform(-2.5, -1, n - 1)
    x0 = np.zeros((n, 3)); v0 = np.zeros((n, 3)); x1 = np.zeros((n, 3)); v1 = np.zeros((n, 3))
    for j in range(1, nb):
        a = rng.uniform(0.5, 5.0)
        th = rng.uniform(0, 2 * np.pi)
        x0[j] = [a * np.cos(th), a * np.sin(th), rng.normal() * 1e-3]
        vc = np.sqrt(MU / a)
        v0[j] = [-vc * np.sin(th), vc * np.cos(th), rng.normal() * 1e-5]
        out, _ = R.drift_one(MU, np.concatenate([x0[j], v0[j]]), h)
        x1[j], v1[j] = out[:3], out[3:]
    # bodies near a planet: Big bodies near an earlier planet (Big-Big pairs,
    # mass order mixed), particles near any planet
    for j in range(2, n):
        if j < nb and rng.uniform() > near:
            continue
        if j >= nb and rng.uniform() > near:
            a = rng.uniform(0.5, 5.0); th = rng.uniform(0, 2 * np.pi)
            x0[j] = [a * np.cos(th), a * np.sin(th), 0.0]
            vc = np.sqrt(MU / a) * rng.uniform(0.9, 1.1)
            v0[j] = [-vc * np.sin(th), vc * np.cos(th), 0.0]
            out, _ = R.drift_one(MU, np.concatenate([x0[j], v0[j]]), h)
            x1[j], v1[j] = out[:3], out[3:]
            continue
        p = int(rng.integers(1, min(j, nb)))
        dv = rng.normal(size=3); dv *= 10 ** rng.uniform(-5, -3) / np.linalg.norm(dv)
        tc = rng.uniform(-1.3, 0.3) * h            # time of closest approach relative to the step end
        e = rng.normal(size=3); e -= e.dot(dv) / dv.dot(dv) * dv; e /= np.linalg.norm(e)
        b = (rphys[p] + rphys[j]) * rng.choice([0.3, 0.9, 1.0, 1.1, 1.7, 2.0, 2.2, 5.0, 50.0])
        if rng.uniform() < 0.3:
            b = rce[j] * rng.uniform(0.2, 1.5)
        d_end = b * e - dv * tc                     # relative offset at the step end
        x1[j] = x1[p] + d_end; v1[j] = v1[p] + dv
        x0[j] = x0[p] + d_end - dv * h; v0[j] = v0[p] + dv
    return dict(m=m, x0=x0, v0=v0, x1=x1, v1=v1, rce=rce, rphys=rphys, nbig=nb, h=h)


def mce_stat(R=None):
    """mce_stat on constructed steps: close encounters, near misses and
    collisions of particles and of Big pairs, both step signs."""
    R = R or Reference()
    rng = np.random.default_rng(20261204)
    out = {}
    shapes = [(8, 200, 4.0, 0.8), (5, 150, -2.0, 0.8), (12, 100, 8.0, 0.6), (3, 300, 1.0, 0.9),
              (8, 0, 4.0, 1.0), (2, 50, -8.0, 1.0)]
    for c, (npl, npart, h, near) in enumerate(shapes):
        q = stat_case(rng, R, npl, npart, h, near)
        time = float(rng.uniform(-100, 100))
        r = R.mce_stat(time, h, q["m"], q["x0"], q["v0"], q["x1"], q["v1"], q["rce"], q["rphys"], q["nbig"])
        for k in ("m", "x0", "v0", "x1", "v1", "rce", "rphys"):
            out[f"{k}{c}"] = q[k]
        out[f"meta{c}"] = np.array([q["nbig"], time, h], float)
        for k, v in r.items():
            out[f"{k}{c}"] = np.asarray(v)
        print(f"mce_stat case {c}: {len(r['iclo'])} encounters, {len(r['ihit'])} near misses/collisions "
              f"({int(np.sum(np.asarray(r['chit']) == 1))} collisions)")
    np.savez_compressed(HERE / "mce_stat.npz", **out)


def nongrav(R=None):
    """mfo_ngf and mfo_pr on random bodies (every parameter, the 88 AU^2 and
    1e-7 cutoffs, massive bodies skipped by mfo_pr), and whole runs with the
    non-gravitational model: RA15 and BS with ngflag 3 and 1PN, BS2 and MVS
    with the cometary terms (ngflag 1).  The MVS run comes last: mdt_mvs
    keeps its cometary accelerations in a saved array."""
    R = R or Reference()
    rng = np.random.default_rng(20261205)
    out = {}
    k = 400
    r = 10 ** rng.uniform(-0.5, 1.6, k)
    d = rng.normal(size=(k, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
    x = np.concatenate([np.zeros((1, 3)), d * r[:, None]])
    v = np.concatenate([np.zeros((1, 3)), rng.normal(size=(k, 3)) * 0.01])
    v[5] = x[5] * 0.001                                   # radial velocity only (A3 normal undefined)
    ngf = np.zeros((k + 1, 5))
    for j in range(1, k + 1):
        for q in range(5):
            if rng.random() < 0.45:
                ngf[j, q] = rng.choice([-1.0, 1.0]) * 10 ** rng.uniform(-10, -6.5)
        if rng.random() < 0.4:
            ngf[j, 3] = 10 ** rng.uniform(-2, -0.5)
    ngf[5] = [1e-8, 0.0, 0.0, 0.0, 0.0]
    m = np.zeros(k + 1); m[0] = MU
    m[1:30] = 1e-9 * MU                                   # massive: mfo_pr skips them
    out["ngf_x"] = x; out["ngf_v"] = v; out["ngf_p"] = ngf; out["ngf_m"] = m
    out["ngf_a"] = R.mfo_ngf(x, v, ngf)
    out["pr_a"] = R.mfo_pr(m, 30, x, v, ngf)
    # HYBRID: cometary terms on particles and Big bodies, half the particles
    # near planet orbits (encounter members go through mdt_hkce)
    s = systems.make_system(3, 300, a_range=(0.5, 40.0), emax=0.3, imax_deg=20.0, h=2.0, near_fraction=0.5)
    systems.add_nongrav(s, seed=51, fraction=0.6, kinds=("a1", "a2", "a3"), big=True)
    for key in ("m", "x", "v", "ngf", "rce", "rphys", "rcrit"):
        out[f"hy_{key}"] = s[key]
    out["hy_meta"] = np.array([s["nbig"], 150, s["h"], s["ngflag"], s["rcen"]], float)
    xo, vo, evi, evd = R.mdt_hy_run(s, 150)
    out["hy_xout"] = xo; out["hy_vout"] = vo; out["hy_evi"] = evi; out["hy_evd"] = evd
    print(f"nongrav run hybrid: ngflag {s['ngflag']}, 150 steps, {len(evi)} event rows")
    runs = [("ra15", 4, 3, 1, 60), ("bs", 2, 3, 1, 30), ("bs2", 3, 1, 0, 30), ("mvs", 1, 1, 0, 300)]
    for name, algor, flag, pn, nsteps in runs:
        s = systems.config_small(seed=30 + algor, nparticles=60)
        kinds = ("a1", "a2", "a3", "yar", "beta") if flag == 3 else ("a1", "a2", "a3")
        systems.add_nongrav(s, seed=40 + algor, fraction=0.6, kinds=kinds, big=(algor != 1))
        assert s["ngflag"] == flag
        for key in ("m", "x", "v", "ngf"):
            out[f"{name}_{key}"] = s[key]
        out[f"{name}_meta"] = np.array([s["nbig"], nsteps, s["h"], flag, pn], float)
        if algor == 4:
            xo, vo, hd, hn = R.ra15_run(s, nsteps, tol=1e-12, pn=pn)
        elif algor in (2, 3):
            xo, vo, hd, hn = R.bs_run(algor, s, nsteps, tol=1e-12, pn=pn)
        else:
            xo, vo, evi, evd = R.mdt_mvs_run(s, nsteps)
            hd, hn = np.zeros(0), 0.0
            out["mvs_xh"], out["mvs_vh"] = R.mvs2h(s, xo, vo)   # the corrector with angf in its kicks
        out[f"{name}_xout"] = xo; out[f"{name}_vout"] = vo; out[f"{name}_hdid"] = np.asarray(hd)
        out[f"{name}_hnext"] = np.array([hn])
        print(f"nongrav run {name}: ngflag {flag}, {int(np.sum(np.any(s['ngf'] != 0, axis=1)))} bodies with "
              f"parameters, {nsteps} steps")
    np.savez_compressed(HERE / "nongrav.npz", **out)


def radau(R=None):
    """mfo_all (gravity, 1PN, oblateness), whole mdt_ra15 runs, mco_mvs2h."""
    R = R or Reference()
    jc = (1e-6 * 0.0046 ** 2, -2e-7 * 0.0046 ** 4, 5e-8 * 0.0046 ** 6)
    cases = {}
    for c, (seed, npl, npart, pn, obl) in enumerate([(300, 8, 40, 0, False), (301, 8, 40, 1, False),
                                                     (302, 3, 25, 1, True), (303, 12, 30, 0, True),
                                                     (304, 0, 10, 1, False)]):
        s = systems.make_system(seed, npart, nplanets=npl)
        s["jcen"] = np.array(jc if obl else (0.0, 0.0, 0.0))
        cases[f"m{c}"] = s["m"]; cases[f"x{c}"] = s["x"]; cases[f"v{c}"] = s["v"]
        cases[f"nbig{c}"] = s["nbig"]; cases[f"pn{c}"] = pn; cases[f"jcen{c}"] = s["jcen"]
        cases[f"a{c}"] = R.mfo_all(s, pn=pn)
    np.savez_compressed(HERE / "mfo_all.npz", **cases)

    runs = {}
    for c, (seed, npart, nsteps, pn, h0, obl) in enumerate([(1, 120, 150, 1, 8.0, False),
                                                            (2, 80, 100, 0, 8.0, False),
                                                            (3, 60, 40, 1, 200.0, False),
                                                            (4, 60, 80, 1, 4.0, True)]):
        s = systems.config_small(seed, npart)
        s["jcen"] = np.array(jc if obl else (0.0, 0.0, 0.0))
        x, v, hd, hn = R.ra15_run(s, nsteps, h0=h0, tol=1e-12, pn=pn)
        for key in ("m", "x", "v", "jcen"):
            runs[f"{key}{c}"] = s[key]
        runs[f"meta{c}"] = np.array([s["nbig"], nsteps, h0, pn, 1e-12])
        runs[f"xout{c}"] = x; runs[f"vout{c}"] = v; runs[f"hdid{c}"] = hd; runs[f"hnext{c}"] = np.array([hn])
    np.savez_compressed(HERE / "ra15_runs.npz", **runs)

    # the symplectic corrector mco_mvs2h on evolved MVS states
    cor = {}
    for c, (seed, npart, jcv) in enumerate([(21, 60, (0.0, 0.0, 0.0)), (22, 40, jc)]):
        s = systems.config_small(seed, npart)
        s["jcen"] = np.array(jcv)
        x, v, _, _ = R.mdt_mvs_run(s, 50)
        xh, vh = R.mvs2h(s, x, v)
        for key in ("m", "jcen"):
            cor[f"{key}{c}"] = s[key]
        cor[f"meta{c}"] = np.array([s["nbig"], s["h"]])
        cor[f"x{c}"] = x; cor[f"v{c}"] = v; cor[f"xh{c}"] = xh; cor[f"vh{c}"] = vh
    np.savez_compressed(HERE / "mvs2h.npz", **cor)


def bs(R=None):
    """Whole runs of algorithm 2 (mdt_bs1) and 3 (mdt_bs2) with mfo_all:
    growing steps, halvings from a too-large first step, 1PN, oblateness."""
    R = R or Reference()
    jc = (1e-6 * 0.0046 ** 2, -2e-7 * 0.0046 ** 4, 5e-8 * 0.0046 ** 6)
    runs = {}
    for c, (algor, seed, npart, nsteps, pn, h0, obl) in enumerate([(2, 11, 60, 40, 0, 4.0, False),
                                                                   (2, 12, 40, 25, 1, 300.0, True),
                                                                   (3, 13, 60, 40, 0, 4.0, True),
                                                                   (3, 14, 40, 25, 1, 300.0, False)]):
        s = systems.config_small(seed, npart)
        s["jcen"] = np.array(jc if obl else (0.0, 0.0, 0.0))
        x, v, hd, hn = R.bs_run(algor, s, nsteps, h0=h0, tol=1e-12, pn=pn)
        for key in ("m", "x", "v", "jcen"):
            runs[f"{key}{c}"] = s[key]
        runs[f"meta{c}"] = np.array([algor, s["nbig"], nsteps, h0, pn, 1e-12])
        runs[f"xout{c}"] = x; runs[f"vout{c}"] = v; runs[f"hdid{c}"] = hd; runs[f"hnext{c}"] = np.array([hn])
    np.savez_compressed(HERE / "bs_runs.npz", **runs)


def hybrid(R=None):
    """Whole HYBRID runs (mdt_hy with BS2 encounter substeps, the driver's
    mce_cent on physical coordinates), collisions not allowed."""
    R = R or Reference()
    runs = {}
    for c, (seed, npart, nsteps, near) in enumerate([(3, 150, 120, 0.5), (6, 80, 80, 0.0), (7, 200, 100, 0.5)]):
        s = systems.make_system(seed, npart, a_range=(0.5, 40.0), emax=0.3, imax_deg=20.0, h=2.0,
                                near_fraction=near)
        x, v, evi, evd = R.mdt_hy_run(s, nsteps)
        for key in ("m", "x", "v", "rce", "rphys", "rcrit"):
            runs[f"{key}{c}"] = s[key]
        runs[f"meta{c}"] = np.array([s["nbig"], nsteps, s["h"], s["rcen"]], float)
        runs[f"xout{c}"] = x; runs[f"vout{c}"] = v; runs[f"evi{c}"] = evi; runs[f"evd{c}"] = evd
    np.savez_compressed(HERE / "hy_runs.npz", **runs)


if __name__ == "__main__":
    threading.stack_size(512 << 20)
    only = sys.argv[1] if len(sys.argv) > 1 else None
    t = threading.Thread(target={"radau": radau, "bs": bs, "hybrid": hybrid, "mce_stat": mce_stat,
                                 "nongrav": nongrav}.get(only, main))
    t.start()
    t.join()
