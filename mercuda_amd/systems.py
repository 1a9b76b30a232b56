"""Seeded synthetic planetary systems: Big bodies plus massless test particles.

The generator draws orbital elements, converts them to heliocentric Cartesian
states, and fills the per-body encounter limits the way the reference driver
does before it uploads them:

* masses are G*M in AU^3/day^2 (the reference multiplies by K2 on input,
  mercury6_2.for:6270);
* close-encounter limit rce = rceh * Hill radius and physical radius rphys from
  the density, as mce_init/mce_hill (mercury6_2.for:1285-1387) with the
  reference's defaults rceh = 1 Hill radius, density 1 g/cm^3 (:6256-6258);
* changeover distance rcrit = max(cefac * Hill, 0.4 * |h| * vmax) (:1381-1387).

A system is a dict with numpy arrays m (n,), x (n,3), v (n,3), rce, rphys,
rcrit (n,), and scalars nbig (Big bodies including the central body), h, rcen,
jcen.  Body 0 is the central body (zero state).
"""
from __future__ import annotations

import math

import numpy as np

K2 = 2.959122082855911e-4           # mercury.inc:41
AU_CM = 1.495978707e13               # mercury.inc:43
MSUN_G = 1.9884158605722262e33       # mercury.inc:44
DEG = math.pi / 180.0


def kepler_E(M, e):
    """Eccentric anomaly for elliptic orbits by Newton iteration (vectorised)."""
    M = np.mod(M, 2 * math.pi)
    E = np.where(e < 0.8, M, math.pi * np.ones_like(M))
    for _ in range(60):
        f = E - e * np.sin(E) - M
        E = E - f / (1.0 - e * np.cos(E))
        if f.size == 0 or np.max(np.abs(f)) < 1e-15:
            break
    return E


def elements_to_xv(gm, a, e, inc, peri, node, M):
    """Heliocentric position/velocity from elliptic elements (angles in rad)."""
    gm, a, e, inc, peri, node, M = np.broadcast_arrays(
        *[np.asarray(q, dtype=np.float64) for q in (gm, a, e, inc, peri, node, M)])
    E = kepler_E(M, e)
    cE, sE = np.cos(E), np.sin(E)
    b = a * np.sqrt(1.0 - e * e)
    n = np.sqrt(gm / a ** 3)
    px, py = a * (cE - e), b * sE
    edot = n / (1.0 - e * cE)
    vx, vy = -a * sE * edot, b * cE * edot
    cw, sw = np.cos(peri), np.sin(peri)
    cn, sn = np.cos(node), np.sin(node)
    ci, si = np.cos(inc), np.sin(inc)
    # rotation R = Rz(node) Rx(inc) Rz(peri)
    r11 = cn * cw - sn * sw * ci
    r12 = -cn * sw - sn * cw * ci
    r21 = sn * cw + cn * sw * ci
    r22 = -sn * sw + cn * cw * ci
    r31 = sw * si
    r32 = cw * si
    x = np.stack([r11 * px + r12 * py, r21 * px + r22 * py, r31 * px + r32 * py], axis=-1)
    v = np.stack([r11 * vx + r12 * vy, r21 * vx + r22 * vy, r31 * vx + r32 * vy], axis=-1)
    return x, v


def osculating_a(gm, x, v):
    r = np.linalg.norm(x, axis=-1)
    v2 = np.sum(v * v, axis=-1)
    return 1.0 / (2.0 / r - v2 / gm), r


def draw_planets(rng, nplanets, mlo=1e-7, mhi=1e-3, alo=0.4, ahi=30.0, emax=0.05,
                 imax_deg=2.0, hill_sep=10.0, max_tries=2000, batch=512):
    """Planet masses (solar) and elements, redrawn until every pair is at least
    hill_sep mutual Hill radii apart (candidate sets are drawn in batches and
    the first acceptable one is kept)."""
    for _ in range(max_tries):
        m = 10.0 ** rng.uniform(math.log10(mlo), math.log10(mhi), (batch, nplanets))
        a = 10.0 ** rng.uniform(math.log10(alo), math.log10(ahi), (batch, nplanets))
        ok = np.ones(batch, bool)
        for i in range(nplanets):
            for j in range(i + 1, nplanets):
                rh = ((m[:, i] + m[:, j]) / 3.0) ** (1.0 / 3.0) * (a[:, i] + a[:, j]) / 2.0
                ok &= np.abs(a[:, i] - a[:, j]) >= hill_sep * rh
        idx = np.flatnonzero(ok)
        if len(idx):
            m, a = m[idx[0]], a[idx[0]]
            order = np.argsort(a)
            m, a = m[order], a[order]
            e = rng.uniform(0.0, emax, nplanets)
            inc = rng.uniform(0.0, imax_deg, nplanets) * DEG
            ang = rng.uniform(0.0, 2 * math.pi, (3, nplanets))
            return m, a, e, inc, ang[0], ang[1], ang[2]
    raise RuntimeError("could not draw a Hill-stable planet set")


def finish_limits(sysd, rceh=1.0, density=1.0, cefac=3.0):
    """rce, rphys, rcrit for every body, as mce_init does (mercury6_2.for:1355-1387)."""
    m, x, v = sysd["m"], sysd["x"], sysd["v"]
    n = len(m)
    rce = np.zeros(n); rphys = np.zeros(n); rcrit = np.zeros(n)
    third = .3333333333333333
    rhocgs = AU_CM ** 3 * K2 / MSUN_G
    hill = np.zeros(n); a = np.zeros(n)
    if n > 1:
        gm = m[0] + m[1:]
        ao, r = osculating_a(gm, x[1:], v[1:])
        ao = np.where(ao <= 0.0, r, ao)
        a[1:] = ao
        hill[1:] = ao * (third * m[1:] / m[0]) ** third
        temp = 2.25 * m[0] / math.pi
        rce[1:] = hill[1:] * rceh
        rphys[1:] = hill[1:] / a[1:] * (temp / (density * rhocgs)) ** third
        amin = np.min(a[1:])
        vmax = math.sqrt(m[0] / amin)
        rcrit[1:] = np.maximum(hill[1:] * cefac, 0.4 * abs(sysd["h"]) * vmax)
    sysd["rce"], sysd["rphys"], sysd["rcrit"] = rce, rphys, rcrit
    return sysd


def make_system(seed, nparticles, *, nplanets=8, a_range=(0.5, 5.0), emax=0.2,
                imax_deg=10.0, h=8.0, rcen=0.005, cefac=3.0, near_fraction=0.0,
                near_width=0.1, near_jitter_deg=5.0, central_mass=1.0):
    """One system: central body + nplanets Big bodies + nparticles massless.

    near_fraction > 0 places that fraction of the particles within
    +-near_width AU of a random planet's orbit, with that planet's angles
    jittered by U[-near_jitter_deg, near_jitter_deg], so encounters fire."""
    rng = np.random.default_rng(seed)
    m0 = central_mass * K2
    pm, pa, pe, pi_, pw, pn, pM = draw_planets(rng, nplanets) if nplanets else ([],) * 7
    pm = np.asarray(pm) * K2
    xp, vp = elements_to_xv(m0 + pm, pa, pe, pi_, pw, pn, pM) if nplanets else (
        np.zeros((0, 3)), np.zeros((0, 3)))
    k = nparticles
    a = rng.uniform(a_range[0], a_range[1], k)
    e = rng.uniform(0.0, emax, k)
    inc = rng.uniform(0.0, imax_deg, k) * DEG
    ang = rng.uniform(0.0, 2 * math.pi, (3, k))
    w, node, M = ang
    if near_fraction > 0 and nplanets > 0 and k > 0:
        near = rng.random(k) < near_fraction
        host = rng.integers(0, nplanets, k)
        jit = lambda: rng.uniform(-near_jitter_deg, near_jitter_deg, k) * DEG
        a = np.where(near, pa[host] + rng.uniform(-near_width, near_width, k), a)
        e = np.where(near, pe[host], e)
        inc = np.where(near, np.abs(pi_[host] + jit()), inc)
        w = np.where(near, pw[host] + jit(), w)
        node = np.where(near, pn[host] + jit(), node)
        M = np.where(near, pM[host] + jit(), M)
    xs, vs = elements_to_xv(np.full(k, m0), a, e, inc, w, node, M)
    n = 1 + nplanets + k
    m = np.zeros(n); m[0] = m0; m[1:1 + nplanets] = pm
    x = np.zeros((n, 3)); v = np.zeros((n, 3))
    x[1:1 + nplanets] = xp; v[1:1 + nplanets] = vp
    x[1 + nplanets:] = xs; v[1 + nplanets:] = vs
    sysd = dict(m=m, x=x, v=v, nbig=1 + nplanets, h=float(h), rcen=float(rcen),
                jcen=np.zeros(3), seed=seed)
    return finish_limits(sysd, cefac=cefac)


# Named workloads (sizes in the docstrings of the functions that use them).
def config_small(seed=1, nparticles=1000):
    """8 planets + 1,000 particles, a U[0.5,5] AU, e U[0,0.2], i U[0,10 deg], h = 8 d."""
    return make_system(seed, nparticles, a_range=(0.5, 5.0), emax=0.2, imax_deg=10.0, h=8.0)


def config_large(seed=2, nparticles=1_000_000):
    """8 planets + 1e5/1e6 particles, a U[0.5,40], e U[0,0.3], i U[0,20 deg], h = 4 d."""
    return make_system(seed, nparticles, a_range=(0.5, 40.0), emax=0.3, imax_deg=20.0, h=4.0)


def config_encounters(seed=3, nparticles=200_000):
    """8 planets + 2e5 particles, half near planet orbits, changeover 3 Hill, h = 2 d."""
    return make_system(seed, nparticles, a_range=(0.5, 40.0), emax=0.3, imax_deg=20.0, h=2.0,
                       near_fraction=0.5, cefac=3.0)


def ensemble(seed=4, nsystems=512, nmin=100, nmax=50_000, h=4.0):
    """nsystems independent systems, 2..12 Big bodies (uniform) and log-uniform
    particle counts in [nmin, nmax]; returns a list of system dicts."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nsystems):
        nbig = int(rng.integers(2, 13))        # Big bodies excluding the central body
        npart = int(round(10.0 ** rng.uniform(math.log10(nmin), math.log10(nmax))))
        out.append(make_system(seed * 100003 + s, npart, nplanets=nbig, a_range=(0.5, 40.0),
                               emax=0.3, imax_deg=20.0, h=h))
    return out


def add_nongrav(sysd, seed, fraction=0.5, kinds=("a1", "a2", "a3", "yar", "beta"), big=False):
    """Non-gravitational parameters (the driver's ngf(5,n): A1, A2, A3 in
    AU/day^2, beta, yar) on a seeded fraction of the particles (and of the Big
    bodies when big is true, except beta, which mfo_pr applies to massless
    bodies only): cometary A values ~1e-9..1e-7, some above the 1e-7 cutoff
    threshold; yar ~1e-12..1e-10; beta ~0.01..0.3.  Sets sysd["ngf"] and
    sysd["ngflag"] (1 cometary and Yarkovsky, 2 radiation, 3 both)."""
    rng = np.random.default_rng(seed)
    n = len(sysd["m"])
    ngf = np.zeros((n, 5))
    first = 1 if big else sysd["nbig"]
    for j in range(first, n):
        if rng.random() > fraction:
            continue
        for k, name in enumerate(("a1", "a2", "a3")):
            if name in kinds and rng.random() < 0.6:
                ngf[j, k] = rng.choice([-1.0, 1.0]) * 10 ** rng.uniform(-9, -6.7)
        if "beta" in kinds and j >= sysd["nbig"] and rng.random() < 0.5:
            ngf[j, 3] = 10 ** rng.uniform(-2, -0.5)
        if "yar" in kinds and rng.random() < 0.5:
            ngf[j, 4] = rng.choice([-1.0, 1.0]) * 10 ** rng.uniform(-12, -10)
    comet = bool(np.any(ngf[:, [0, 1, 2, 4]] != 0))
    rad = bool(np.any(ngf[:, 3] != 0))
    sysd["ngf"] = ngf
    sysd["ngflag"] = (1 if comet else 0) + (2 if rad else 0)
    return sysd


def config_extra_forces(seed=1729, nparticles=100_000):
    """The reference's published extra-force case (docs/benchmarks.md:36-44,
    75-80): 8 planets + 1e5 massless particles at perihelion with random
    azimuths, a U[3.2,3.8] AU, e U[0.01,0.05] (coplanar here), each with
    beta = 1e-4 (radiation pressure and PR drag) and a Yarkovsky coefficient
    of 1e-12 AU/day^2; first step 1 d.  The planet set is this module's
    Hill-stable draw with no planet within 1 AU of the particle belt (the
    reference's eight-planet system is synthetic and unpublished)."""
    rng = np.random.default_rng(seed)
    for _ in range(1000):
        base = make_system(int(rng.integers(1 << 30)), 0, h=1.0)
        ap, _ = osculating_a(base["m"][0] + base["m"][1:], base["x"][1:], base["v"][1:])
        if np.all((ap < 2.2) | (ap > 4.8)):
            break
    k = nparticles
    a = rng.uniform(3.2, 3.8, k)
    e = rng.uniform(0.01, 0.05, k)
    w = rng.uniform(0.0, 2 * math.pi, k)
    zero = np.zeros(k)
    xs, vs = elements_to_xv(np.full(k, base["m"][0]), a, e, zero, w, zero, zero)
    nb = base["nbig"]
    s = dict(m=np.concatenate([base["m"], np.zeros(k)]), x=np.concatenate([base["x"], xs]),
             v=np.concatenate([base["v"], vs]), nbig=nb, h=1.0, rcen=base["rcen"], jcen=np.zeros(3), seed=seed)
    s = finish_limits(s)
    ngf = np.zeros((nb + k, 5))
    ngf[nb:, 3] = 1e-4
    ngf[nb:, 4] = 1e-12
    s["ngf"] = ngf
    s["ngflag"] = 3
    return s
