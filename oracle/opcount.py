"""Algorithmic FP64 operation counts of one MVS particle-step and of one RA15
particle-sequence -- TEST/MEASUREMENT INFRASTRUCTURE ONLY.

The particle side of mdt_mvs (mercury6_2.for:4057-4097: kick, drift_one,
mce_stat's particle box and pair screen, mfo_mvs small-body force, kick) and
the driver's mce_cent (mercury6_2.for:1059-1185) are restated here over a float
type that counts every IEEE +, -, *, / and sqrt it executes (each counts 1;
there are no fused multiply-adds in the reference build, -ffp-contract=off;
comparisons, min/max, sign transfers and conversions count 0; libm calls count
as one operation).  Python floats are IEEE doubles with correctly rounded
arithmetic, so the restatement reproduces the oracle bit for bit -- checked by
``verify`` -- and its counts are the counts of the computation the HIP kernels
perform.

    python -m oracle.opcount        # prints the per-particle-step averages
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

OPS = [0]


class F:
    __slots__ = ("v",)

    def __init__(self, v):
        self.v = float(v)

    @staticmethod
    def _v(o):
        return o.v if isinstance(o, F) else float(o)

    def __add__(self, o): OPS[0] += 1; return F(self.v + F._v(o))
    def __radd__(self, o): OPS[0] += 1; return F(F._v(o) + self.v)
    def __sub__(self, o): OPS[0] += 1; return F(self.v - F._v(o))
    def __rsub__(self, o): OPS[0] += 1; return F(F._v(o) - self.v)
    def __mul__(self, o): OPS[0] += 1; return F(self.v * F._v(o))
    def __rmul__(self, o): OPS[0] += 1; return F(F._v(o) * self.v)
    def __truediv__(self, o): OPS[0] += 1; return F(self.v / F._v(o))
    def __rtruediv__(self, o): OPS[0] += 1; return F(F._v(o) / self.v)
    def __neg__(self): return F(-self.v)
    def __lt__(self, o): return self.v < F._v(o)
    def __le__(self, o): return self.v <= F._v(o)
    def __gt__(self, o): return self.v > F._v(o)
    def __ge__(self, o): return self.v >= F._v(o)
    def __eq__(self, o): return self.v == F._v(o)


def sqrt(x): OPS[0] += 1; return F(math.sqrt(F._v(x)))
def fn1(f, x): OPS[0] += 1; return F(f(F._v(x)))
def fmin(a, b): return a if F._v(a) <= F._v(b) else b
def fmax(a, b): return a if F._v(a) >= F._v(b) else b
def fabs(a): return F(abs(F._v(a)))
def sign(a, b): return F(math.copysign(abs(F._v(a)), F._v(b)))


DANBYB = 1.0e-13
TWOPI = 2.0 * 3.141592653589793


def stumpff(x):
    n = 0
    while fabs(x) >= 0.1:
        n += 1; x = x * .25
    x2 = x * x; x3 = x * x2; x4 = x2 * x2; x5 = x2 * x3; x6 = x3 * x3
    c2 = (1.147074559772972e-11 * x6 - 2.087675698786810e-9 * x5 + 2.755731922398589e-7 * x4
          - 2.480158730158730e-5 * x3 + 1.388888888888889e-3 * x2 - 4.166666666666667e-2 * x + .5)
    c3 = (7.647163731819816e-13 * x6 - 1.605904383682161e-10 * x5 + 2.505210838544172e-8 * x4
          - 2.755731922398589e-6 * x3 + 1.984126984126984e-4 * x2 - 8.333333333333333e-3 * x
          + 1.666666666666667e-1)
    c1 = 1.0 - x * c3
    c0 = 1.0 - x * c2
    for _ in range(n):
        c3 = (c2 + c0 * c3) * .25
        c2 = c1 * c1 * .5
        c1 = c0 * c1
        c0 = 2.0 * c0 * c0 - 1.0
        x = x * 4.0
    return c0, c1, c2, c3


def kepu_fchk(dt, r0, mu, alpha, u, s):
    c0, c1, c2, c3 = stumpff(s * s * alpha)
    c1 = c1 * s; c2 = c2 * s * s; c3 = c3 * s * s * s
    return r0 * c1 + u * c2 + mu * c3 - dt


def p3solve(dt, r0, mu, alpha, u):
    denom = (mu - alpha * r0) / 6.0
    a2 = 0.5 * u / denom; a1 = r0 / denom; a0 = -dt / denom
    q = (a1 - a2 * a2 / 3.0) / 3.0
    r = (a1 * a2 - 3.0 * a0) / 6.0 - (a2 * a2 * a2) / 27.0
    sq2 = (q * q * q) + (r * r)
    if sq2 >= 0.0:
        sq = sqrt(sq2)
        third = 1.0 / 3.0
        p1 = -fn1(lambda z: (-z) ** third, r + sq) if (r + sq) <= 0.0 else fn1(lambda z: z ** third, r + sq)
        p2 = -fn1(lambda z: (-z) ** third, r - sq) if (r - sq) <= 0.0 else fn1(lambda z: z ** third, r - sq)
        return p1 + p2 - a2 / 3.0, 0
    return F(0.0), 1


def kepu_guess(dt, r0, mu, alpha, u):
    if alpha > 0.0:
        if dt / r0 <= 0.4:
            return dt / r0 - (dt * dt * u) / (2.0 * r0 * r0 * r0)
        a = mu / alpha
        en = sqrt(mu / (a * a * a))
        ec = 1.0 - r0 / a
        es = u / (en * a * a)
        e = sqrt(ec * ec + es * es)
        y = en * dt - es
        y = fn1(lambda z: math.fmod(z, TWOPI), y) if y > 0.0 else fn1(lambda z: math.fmod(z, TWOPI), y) + TWOPI
        cy = fn1(math.cos, y)
        sy = -sqrt(1.0 - cy * cy) if y > 3.141592653589793 else sqrt(1.0 - cy * cy)
        sigma = sign(1.0, es * cy + ec * sy)
        x = y + sigma * .85 * e
        return x / sqrt(alpha)
    s, fl = p3solve(dt, r0, mu, alpha, u)
    return dt / r0 if fl else s


def kepu_new(s, dt, r0, mu, alpha, u):
    for _ in range(7):
        s2 = s * s
        c0, c1, c2, c3 = stumpff(s2 * alpha)
        c1 = c1 * s; c2 = c2 * s2; c3 = c3 * s * s2
        f = r0 * c1 + u * c2 + mu * c3 - dt
        fp = r0 * c0 + u * c1 + mu * c2
        fpp = (mu - r0 * alpha) * c1 + u * c0
        fppp = (mu - r0 * alpha) * c0 - u * alpha * c1
        ds = -(f / fp)
        ds = -(f / (fp + .5 * ds * fpp))
        ds = -(f / (fp + .5 * ds * fpp + ds * ds * fppp * .1666666666666667))
        s = s + ds
        fdt = f / dt
        if fdt * fdt < DANBYB * DANBYB:
            return s, fp, c1, c2, c3, 0
    return s, fp, c1, c2, c3, 1


def kepu_lag(s, dt, r0, mu, alpha, u):
    ln = 5.0
    for _ in range(401):
        c0, c1, c2, c3 = stumpff(s * s * alpha)
        c1 = c1 * s; c2 = c2 * s * s; c3 = c3 * s * s * s
        f = r0 * c1 + u * c2 + mu * c3 - dt
        fp = r0 * c0 + u * c1 + mu * c2
        fpp = (-40.0 * alpha + mu) * c1 + u * c0
        ds = -(ln * f / (fp + sign(1.0, fp) * sqrt(fabs((ln - 1.0) * (ln - 1.0) * fp * fp
                                                        - (ln - 1.0) * ln * f * fpp))))
        s = s + ds
        fdt = f / dt
        if fdt * fdt < DANBYB * DANBYB:
            return s, fp, c1, c2, c3, 0
    return s, fp, c1, c2, c3, 2


def kepu(dt, r0, mu, alpha, u):
    s = kepu_guess(dt, r0, mu, alpha, u)
    st = s
    s, fp, c1, c2, c3, fl = kepu_new(s, dt, r0, mu, alpha, u)
    if fl:
        fo = kepu_fchk(dt, r0, mu, alpha, u, st)
        fn = kepu_fchk(dt, r0, mu, alpha, u, s)
        if fabs(fo) < fabs(fn):
            s = st
        s, fp, c1, c2, c3, fl = kepu_lag(s, dt, r0, mu, alpha, u)
    return fp, c1, c2, c3, fl


def kepmd(dm, es, ec):
    A0, A1, A2, A3, A4 = 39916800.0, 6652800.0, 332640.0, 7920.0, 110.0
    fac1 = 1.0 / (1.0 - ec)
    q = fac1 * dm
    fac2 = es * es * fac1 - ec / 3.0
    x = q * (1.0 - 0.5 * fac1 * q * (es - q * fac2))
    y = x * x
    s = x * (A0 - y * (A1 - y * (A2 - y * (A3 - y * (A4 - y))))) / A0
    c = sqrt(1.0 - s * s)
    f = x - ec * s + es * (1.0 - c) - dm
    fp = 1. - ec * c + es * s
    fpp = ec * s + es * c
    fppp = ec * c - es * s
    dx = -(f / fp)
    dx = -(f / (fp + 0.5 * dx * fpp))
    dx = -(f / (fp + 0.5 * dx * fpp + 0.16666666666666666 * dx * dx * fppp))
    x = x + dx
    y = x * x
    s = x * (A0 - y * (A1 - y * (A2 - y * (A3 - y * (A4 - y))))) / A0
    c = sqrt(1.0 - s * s)
    return x, s, c


def drift_dan(mu, st, dt0):
    x0, y0, z0, vx0, vy0, vz0 = st
    dt = dt0
    r0 = sqrt(x0 * x0 + y0 * y0 + z0 * z0)
    v0s = vx0 * vx0 + vy0 * vy0 + vz0 * vz0
    u = x0 * vx0 + y0 * vy0 + z0 * vz0
    alpha = 2.0 * mu / r0 - v0s
    if alpha > 0.0:
        a = mu / alpha
        asq = a * a
        en = sqrt(mu / (a * asq))
        ec = 1.0 - r0 / a
        es = u / (en * asq)
        esq = ec * ec + es * es
        ten = dt * en
        k = int((dt * en / TWOPI).v)
        dm = ten - float(k) * TWOPI
        dt = dm / en
        if not (dm * dm > 0.16 or esq > 0.36) and esq * dm * dm < 0.0016:
            xkep, s, c = kepmd(dm, es, ec)
            fchk = xkep - ec * s + es * (1.0 - c) - dm
            if fchk * fchk > DANBYB:
                return st, 1, "kepmd"
            fp = 1.0 - ec * c + es * s
            f = (a / r0) * (c - 1.0) + 1.0
            g = dt + (s - xkep) / en
            fdot = -((a / (r0 * fp)) * en * s)
            gdot = (c - 1.) / fp + 1.
            return (x0 * f + vx0 * g, y0 * f + vy0 * g, z0 * f + vz0 * g,
                    x0 * fdot + vx0 * gdot, y0 * fdot + vy0 * gdot, z0 * fdot + vz0 * gdot), 0, "kepmd"
    fp, c1, c2, c3, fl = kepu(dt, r0, mu, alpha, u)
    if fl == 0:
        f = 1.0 - (mu / r0) * c2
        g = dt - mu * c3
        fdot = -((mu / (fp * r0)) * c1)
        gdot = 1.0 - (mu / fp) * c2
        st = (x0 * f + vx0 * g, y0 * f + vy0 * g, z0 * f + vz0 * g,
              x0 * fdot + vx0 * gdot, y0 * fdot + vy0 * gdot, z0 * fdot + vz0 * gdot)
    return st, fl, "kepu"


def drift_one(mu, st, dt):
    st, fl, br = drift_dan(mu, st, dt)
    if fl:
        br = "retry"
        for _ in range(10):
            st, fl, _ = drift_dan(mu, st, dt / 10.0)
            if fl:
                break
    return st, br


def box_axis(h, x0, v0, x1, v1):
    lo, hi = fmin(x0, x1), fmax(x0, x1)
    if (v0 < 0.0 and v1 > 0.0) or (v0 > 0.0 and v1 < 0.0):
        t = (v0 * x1 - v1 * x0 - .5 * h * v0 * v1) / (v0 - v1)
        lo, hi = fmin(lo, t), fmax(hi, t)
    return lo, hi


def particle_step(mu, h, state, acc, planets, rcen, rce_j=0.0, rphys_j=0.0, mj=0.0):
    """One MVS particle-step with the driver's screens.  planets: per-step dict
    with a0tp (3), X0, V0, X1, V1 (np x 3), m, box (np x 4 padded), rce, rphys."""
    hby2 = F(.5) * h
    x, y, z, vx, vy, vz = state
    ax, ay, az = acc
    xs = (x, y, z, vx, vy, vz)
    vx = vx + hby2 * ax; vy = vy + hby2 * ay; vz = vz + hby2 * az
    k0 = (x, y, z, vx, vy, vz)
    st, br = drift_one(mu, k0, h)
    x, y, z, vx, vy, vz = st
    # mce_stat: particle box, pad, then per planet the overlap test and detail
    xl, xh = box_axis(h, k0[0], k0[3], x, vx)
    yl, yh = box_axis(h, k0[1], k0[4], y, vy)
    pad = F(rce_j) * 1.2
    xl = xl - pad; xh = xh + pad; yl = yl - pad; yh = yh + pad
    nd = 0
    for p in range(len(planets["m"])):
        b = planets["box"][p]
        if b[1] >= xl.v and xh.v >= b[0] and b[3] >= yl.v and yh.v >= b[2]:
            nd += 1
            X0, V0, X1, V1 = (planets[k][p] for k in ("X0", "V0", "X1", "V1"))
            dx0 = X0[0] - k0[0]; dy0 = X0[1] - k0[1]; dz0 = X0[2] - k0[2]
            du0 = V0[0] - k0[3]; dv0 = V0[1] - k0[4]; dw0 = V0[2] - k0[5]
            d0t = (dx0 * du0 + dy0 * dv0 + dz0 * dw0) * 2.0
            dx1 = X1[0] - x; dy1 = X1[1] - y; dz1 = X1[2] - z
            du1 = V1[0] - vx; dv1 = V1[1] - vy; dw1 = V1[2] - vz
            d1t = (dx1 * du1 + dy1 * dv1 + dz1 * dw1) * 2.0
            d0 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0
            d1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1
            # mce_min
            if d0t * h > 0.0 or d1t * h < 0.0:
                pass
            else:
                temp = 6.0 * (d0 - d1)
                a = temp + 3.0 * h * (d0t + d1t)
                bb = temp + 2.0 * h * (d0t + 2.0 * d1t)
                c = h * d1t
                temp = -.5 * (bb + sign(sqrt(fmax(bb * bb - 4.0 * a * c, F(0.0))), bb))
                tau = F(0.0) if temp == 0.0 else c / temp
                tau = fmax(fmin(tau, F(0.0)), F(-1.0))
                temp = 1.0 + tau
                _ = tau * tau * ((3.0 + 2.0 * tau) * d0 + temp * h * d0t) \
                    + temp * temp * ((1.0 - 2.0 * tau) * d1 + tau * h * d1t)
            d2ce = fmax(F(planets["rce"][p]), F(rce_j)); d2hit = F(planets["rphys"][p]) + rphys_j
            d2ce = d2ce * d2ce; d2hit = d2hit * d2hit; _ = d2hit * 4.0
    # mfo_mvs small-body force
    a3x = F(0.0); a3y = F(0.0); a3z = F(0.0)
    for p in range(len(planets["m"])):
        X1 = planets["X1"][p]
        dx = x - X1[0]; dy = y - X1[1]; dz = z - X1[2]
        s2 = dx * dx + dy * dy + dz * dz
        s_1 = 1.0 / sqrt(s2)
        s_3 = s_1 * s_1 * s_1
        faci = planets["m"][p] * s_3
        a3x = a3x - faci * dx; a3y = a3y - faci * dy; a3z = a3z - faci * dz
    a0 = planets["a0tp"]
    ax = a0[0] + a3x; ay = a0[1] + a3y; az = a0[2] + a3z
    vx = vx + hby2 * ax; vy = vy + hby2 * ay; vz = vz + hby2 * az
    # mce_cent gate (the impact branch is rare and not counted separately)
    rr0 = xs[0] * xs[0] + xs[1] * xs[1] + xs[2] * xs[2]
    rr1 = x * x + y * y + z * z
    rv0 = xs[3] * xs[0] + xs[4] * xs[1] + xs[5] * xs[2]
    rv1 = vx * x + vy * y + vz * z
    if (rv0 * h <= 0.0 and rv1 * h >= 0.0) or fmin(rr0, rr1) <= rcen * rcen:
        hx = xs[1] * xs[5] - xs[2] * xs[4]; hy = xs[2] * xs[3] - xs[0] * xs[5]; hz = xs[0] * xs[4] - xs[1] * xs[3]
        v2 = xs[3] * xs[3] + xs[4] * xs[4] + xs[5] * xs[5]
        h2 = hx * hx + hy * hy + hz * hz
        pp = h2 / (mu + mj)
        r0 = sqrt(rr0)
        temp = 1.0 + pp * (v2 / (mu + mj) - 2.0 / r0)
        e = sqrt(fmax(temp, F(0.0)))
        _ = pp / (1.0 + e)
    return (x, y, z, vx, vy, vz), (ax, ay, az), br, nd


def tables_from_oracle(sysd, nsteps):
    """Per-step planet data for the particle steps, from the oracle's own whole-
    system steps (planets only)."""
    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    from oracle.bindings import Oracle
    O = Oracle()
    nb = sysd["nbig"]
    pl = {k: sysd[k][:nb].copy() for k in ("m", "x", "v", "rce", "rphys")}
    pl.update(nbig=nb, h=sysd["h"], rcen=sysd["rcen"], jcen=np.zeros(3))
    out = []
    h = sysd["h"]
    x = pl["x"].copy(); v = pl["v"].copy()
    a0_start = O.mfo_mvs(pl["m"], x, nb)
    for _ in range(nsteps):
        # replicate one planet step via the oracle on the planets alone
        sub = dict(pl, x=x, v=v)
        a_now = O.mfo_mvs(pl["m"], x, nb)
        vk = v + 0.5 * h * a_now
        x1, v1k2, _, _ = O.mvs_run(sub, 1, screen=False)
        a_end = O.mfo_mvs(pl["m"], x1, nb)
        v1 = v1k2 - 0.5 * h * a_end      # velocity after the drift (for screening only)
        # indirect term at the step end (planets' a0tp): acceleration of a
        # particle placed far away minus its direct terms is not needed; use mfo
        m = pl["m"]
        a0tp = -sum(m[k] * x1[k] / np.linalg.norm(x1[k]) ** 3 for k in range(1, nb))
        box = []
        for p in range(1, nb):
            pad = pl["rce"][p] * 1.2
            bx = [min(x[p][0], x1[p][0]) - pad, max(x[p][0], x1[p][0]) + pad,
                  min(x[p][1], x1[p][1]) - pad, max(x[p][1], x1[p][1]) + pad]
            box.append(bx)
        out.append(dict(a0tp=a0tp, X0=x[1:], V0=vk[1:], X1=x1[1:], V1=v1[1:], m=m[1:], box=box,
                        rce=pl["rce"][1:], rphys=pl["rphys"][1:]))
        x, v = x1, v1k2
    return out, a0_start


def measure(sysd, nsample=1500, nsteps=12, seed=0):
    """Average counted operations per particle-step over a random particle sample
    of the system, split by drift path."""
    tabs, _ = tables_from_oracle(sysd, nsteps)
    rng = np.random.default_rng(seed)
    nb = sysd["nbig"]
    idx = rng.choice(np.arange(nb, len(sysd["m"])), min(nsample, len(sysd["m"]) - nb), replace=False)
    mu = F(sysd["m"][0]); h = F(sysd["h"])
    counts = {"kepmd": [0, 0], "kepu": [0, 0], "retry": [0, 0]}
    for j in idx:
        st = tuple(F(q) for q in np.concatenate([sysd["x"][j], sysd["v"][j]]))
        # start acceleration from the start positions (reset step)
        x0 = sysd["x"][:nb]
        a = -sum(sysd["m"][k] * x0[k] / np.linalg.norm(x0[k]) ** 3 for k in range(1, nb))
        a = a - sum(sysd["m"][k] * (sysd["x"][j] - x0[k]) / np.linalg.norm(sysd["x"][j] - x0[k]) ** 3
                    for k in range(1, nb))
        acc = tuple(F(q) for q in a)
        for s in range(nsteps):
            OPS[0] = 0
            st, acc, br, _ = particle_step(mu, h, st, acc, tabs[s], F(sysd["rcen"]))
            counts[br][0] += OPS[0]
            counts[br][1] += 1
    tot = sum(c[0] for c in counts.values()); n = sum(c[1] for c in counts.values())
    res = {k: (c[0] / c[1] if c[1] else 0.0, c[1]) for k, c in counts.items()}
    res["all"] = (tot / n, n)
    return res


def verify(sysd, nsample=300, seed=1):
    """The counted restatement's drift equals the oracle's bit for bit."""
    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    from oracle.bindings import Oracle
    O = Oracle()
    rng = np.random.default_rng(seed)
    nb = sysd["nbig"]
    idx = rng.choice(np.arange(nb, len(sysd["m"])), nsample, replace=False)
    for j in idx:
        xv = np.concatenate([sysd["x"][j], sysd["v"][j]])
        for dt in (sysd["h"], 40.0 * sysd["h"], -3.0 * sysd["h"]):
            want, _ = O.drift_one(sysd["m"][0], xv, dt)
            got, _ = drift_one(F(sysd["m"][0]), tuple(F(q) for q in xv), F(dt))
            if not np.array_equal(np.array([g.v for g in got]), want):
                return False
    return True


# ---------------------------------------------------------------------------
# RA15: one sequence of mdt_ra15 (mercury6_2.for:4133-4505) with mfo_all's
# gravity and central 1PN (mfo_grav, mercury6_2.for:4593-4707), no oblateness
# or non-gravitational terms.  The whole system is restated; the particles'
# rows are F values, so the count covers the operations on each particle's
# own data flow: its predictors, its force (the planets' direct terms, the
# Sun's term with 1PN, plus the indirect sum the planets' table supplies), its
# corrector and its end-of-sequence update and b/e prediction.  The planets'
# work (plain floats) is per system, not per particle, and is not counted.
# A massless particle's zero-mass terms on the planets and in the indirect sum
# change no bit of the reference's result (x - 0 and x + 0 for nonzero x), so
# they are left out, as the device kernels leave them out.

RA_H = (0.0, .0562625605369221, .1802406917368924, .3526247171131696,
        .5471536263305554, .7342101772154105, .8853209468390958, .9775206135612875)   # :4162-4165
RA_XC = (.5, .1666666666666667, .08333333333333333, .05, .03333333333333333,
         .02380952380952381, .01785714285714286, .01388888888888889)                   # :4166-4168
RA_VC = (.5, .3333333333333333, .25, .2, .1666666666666667, .1428571428571429, .125)   # :4169-4170


def ra_constants():
    """r, c, d of mdt_ra15 (mercury6_2.for:4176-4199), 0-based."""
    h = RA_H
    r = [1.0 / (h[j] - h[k]) for j in range(1, 8) for k in range(j)]
    c = [0.0] * 21; d = [0.0] * 21
    c[0] = -h[1]; d[0] = h[1]
    n = 1
    for j in range(3, 8):
        n += 1
        c[n - 1] = -h[j - 1] * c[n - j + 1]
        d[n - 1] = h[1] * d[n - j + 1]
        for k in range(3, j):
            n += 1
            c[n - 1] = c[n - j] - h[j - 1] * c[n - j + 1]
            d[n - 1] = d[n - j] + h[k - 1] * d[n - j + 1]
        n += 1
        c[n - 1] = c[n - j] - h[j - 1]
        d[n - 1] = d[n - j] + h[j - 1]
    return r, c, d


_CL = 299792458.0 * 86400.0 / (1.495978707e13 * .01)
LIGHT2 = _CL * _CL


def grav_pn(m, nbig, x, v):
    """mfo_grav with 1PN for every body j >= 1: x, v lists of [3] rows (floats
    for the Big bodies, F for the particles)."""
    nbod = len(x)
    a = [[0.0, 0.0, 0.0] for _ in range(nbod)]
    ri = [None] * nbod; r3 = [None] * nbod
    for i in range(1, nbod):
        s2 = x[i][0] * x[i][0] + x[i][1] * x[i][1] + x[i][2] * x[i][2]
        s_1 = 1.0 / (sqrt(s2) if isinstance(s2, F) else math.sqrt(s2))
        ri[i] = s_1
        r3[i] = s_1 * s_1 * s_1
    sx = sy = sz = 0.0
    for i in range(1, nbig):                  # massless bodies add zero terms
        tmp1 = m[i] * r3[i]
        sx = sx - tmp1 * x[i][0]; sy = sy - tmp1 * x[i][1]; sz = sz - tmp1 * x[i][2]
    for i in range(1, nbig):
        for j in range(i + 1, nbod):
            dx = x[j][0] - x[i][0]; dy = x[j][1] - x[i][1]; dz = x[j][2] - x[i][2]
            s2 = dx * dx + dy * dy + dz * dz
            s_1 = 1.0 / (sqrt(s2) if isinstance(s2, F) else math.sqrt(s2))
            s_3 = s_1 * s_1 * s_1
            tmp1 = s_3 * m[i]
            a[j][0] = a[j][0] - tmp1 * dx; a[j][1] = a[j][1] - tmp1 * dy; a[j][2] = a[j][2] - tmp1 * dz
            if j < nbig:
                tmp2 = s_3 * m[j]
                a[i][0] = a[i][0] + tmp2 * dx; a[i][1] = a[i][1] + tmp2 * dy; a[i][2] = a[i][2] + tmp2 * dz
    s = (sx, sy, sz)
    for i in range(1, nbod):
        xi, vi = x[i], v[i]
        tmp1 = m[0] * r3[i]
        rv = xi[0] * vi[0] + xi[1] * vi[1] + xi[2] * vi[2]
        v2 = vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]
        f = tmp1 / LIGHT2
        radial = 4.0 * m[0] * ri[i] - v2
        for c in range(3):
            a[i][c] = a[i][c] + s[c] - tmp1 * xi[c] + f * (radial * xi[c] + 4.0 * rv * vi[c])
    return a


class Ra15State:
    """mdt_ra15's saved b, e and dtflag for the bodies j >= 1 (rows [j][c][7])."""

    def __init__(self, nbod):
        self.dtflag = 0
        self.b = [[[0.0] * 7 for _ in range(3)] for _ in range(nbod)]
        self.e = [[[0.0] * 7 for _ in range(3)] for _ in range(nbod)]


def ra15_sequence(m, nbig, x1, v1, t, tol, st):
    """One mdt_ra15 sequence with 1PN from trial size t (x1, v1 updated in
    place); returns (tdid, tnext, niter)."""
    h = RA_H
    r, c, d = ra_constants()
    nbod = len(x1)
    if st.dtflag == 0:
        st.dtflag = 1
    b, e = st.b, st.e
    while True:
        if st.dtflag == 1:
            niter = 6
            for j in range(1, nbod):
                for cc in range(3):
                    b[j][cc] = [0.0] * 7; e[j][cc] = [0.0] * 7
        else:
            niter = 2
        a1 = grav_pn(m, nbig, x1, v1)
        g = [[None] * 3 for _ in range(nbod)]
        for j in range(1, nbod):
            for cc in range(3):
                bk = b[j][cc]
                g[j][cc] = [bk[6] * d[15] + bk[5] * d[10] + bk[4] * d[6] + bk[3] * d[3] + bk[2] * d[1] + bk[1] * d[0] + bk[0],
                            bk[6] * d[16] + bk[5] * d[11] + bk[4] * d[7] + bk[3] * d[4] + bk[2] * d[2] + bk[1],
                            bk[6] * d[17] + bk[5] * d[12] + bk[4] * d[8] + bk[3] * d[5] + bk[2],
                            bk[6] * d[18] + bk[5] * d[13] + bk[4] * d[9] + bk[3],
                            bk[6] * d[19] + bk[5] * d[14] + bk[4],
                            bk[6] * d[20] + bk[5],
                            bk[6]]
        for _n in range(niter):
            for j in range(2, 9):
                hj = h[j - 1]
                sv = [0.0] * 9
                sv[0] = t * hj
                sv[1] = sv[0] * sv[0] * .5
                sv[2] = sv[1] * hj * .3333333333333333
                sv[3] = sv[2] * hj * .5
                sv[4] = sv[3] * hj * .6
                sv[5] = sv[4] * hj * .6666666666666667
                sv[6] = sv[5] * hj * .7142857142857143
                sv[7] = sv[6] * hj * .75
                sv[8] = sv[7] * hj * .7777777777777778
                x = [[0.0] * 3 for _ in range(nbod)]
                v = [[0.0] * 3 for _ in range(nbod)]
                for q in range(1, nbod):
                    for cc in range(3):
                        bk = b[q][cc]
                        x[q][cc] = (sv[8] * bk[6] + sv[7] * bk[5] + sv[6] * bk[4] + sv[5] * bk[3] + sv[4] * bk[2]
                                    + sv[3] * bk[1] + sv[2] * bk[0] + sv[1] * a1[q][cc] + sv[0] * v1[q][cc] + x1[q][cc])
                sw = [0.0] * 8
                sw[0] = t * hj
                sw[1] = sw[0] * hj * .5
                sw[2] = sw[1] * hj * .6666666666666667
                sw[3] = sw[2] * hj * .75
                sw[4] = sw[3] * hj * .8
                sw[5] = sw[4] * hj * .8333333333333333
                sw[6] = sw[5] * hj * .8571428571428571
                sw[7] = sw[6] * hj * .875
                for q in range(1, nbod):
                    for cc in range(3):
                        bk = b[q][cc]
                        v[q][cc] = (sw[7] * bk[6] + sw[6] * bk[5] + sw[5] * bk[4] + sw[4] * bk[3] + sw[3] * bk[2]
                                    + sw[2] * bk[1] + sw[1] * bk[0] + sw[0] * a1[q][cc] + v1[q][cc])
                a = grav_pn(m, nbig, x, v)
                for q in range(1, nbod):
                    for cc in range(3):
                        bk, gk = b[q][cc], g[q][cc]
                        ak, a1k = a[q][cc], a1[q][cc]
                        if j == 2:
                            temp = gk[0]
                            gk[0] = (ak - a1k) * r[0]
                            bk[0] = bk[0] + gk[0] - temp
                            continue
                        gkk = ak - a1k
                        if j == 3:
                            temp = gk[1]; gk[1] = (gkk * r[1] - gk[0]) * r[2]; temp = gk[1] - temp
                            bk[0] = bk[0] + temp * c[0]; bk[1] = bk[1] + temp
                        elif j == 4:
                            temp = gk[2]; gk[2] = ((gkk * r[3] - gk[0]) * r[4] - gk[1]) * r[5]; temp = gk[2] - temp
                            bk[0] = bk[0] + temp * c[1]; bk[1] = bk[1] + temp * c[2]; bk[2] = bk[2] + temp
                        elif j == 5:
                            temp = gk[3]
                            gk[3] = (((gkk * r[6] - gk[0]) * r[7] - gk[1]) * r[8] - gk[2]) * r[9]
                            temp = gk[3] - temp
                            bk[0] = bk[0] + temp * c[3]; bk[1] = bk[1] + temp * c[4]; bk[2] = bk[2] + temp * c[5]
                            bk[3] = bk[3] + temp
                        elif j == 6:
                            temp = gk[4]
                            gk[4] = ((((gkk * r[10] - gk[0]) * r[11] - gk[1]) * r[12] - gk[2]) * r[13] - gk[3]) * r[14]
                            temp = gk[4] - temp
                            bk[0] = bk[0] + temp * c[6]; bk[1] = bk[1] + temp * c[7]; bk[2] = bk[2] + temp * c[8]
                            bk[3] = bk[3] + temp * c[9]; bk[4] = bk[4] + temp
                        elif j == 7:
                            temp = gk[5]
                            gk[5] = (((((gkk * r[15] - gk[0]) * r[16] - gk[1]) * r[17] - gk[2]) * r[18] - gk[3]) * r[19]
                                     - gk[4]) * r[20]
                            temp = gk[5] - temp
                            bk[0] = bk[0] + temp * c[10]; bk[1] = bk[1] + temp * c[11]; bk[2] = bk[2] + temp * c[12]
                            bk[3] = bk[3] + temp * c[13]; bk[4] = bk[4] + temp * c[14]; bk[5] = bk[5] + temp
                        else:
                            temp = gk[6]
                            gk[6] = ((((((gkk * r[21] - gk[0]) * r[22] - gk[1]) * r[23] - gk[2]) * r[24] - gk[3]) * r[25]
                                      - gk[4]) * r[26] - gk[5]) * r[27]
                            temp = gk[6] - temp
                            bk[0] = bk[0] + temp * c[15]; bk[1] = bk[1] + temp * c[16]; bk[2] = bk[2] + temp * c[17]
                            bk[3] = bk[3] + temp * c[18]; bk[4] = bk[4] + temp * c[19]; bk[5] = bk[5] + temp * c[20]
                            bk[6] = bk[6] + temp
        # the step-size law (per system, floats)
        temp = 0.0
        for q in range(1, nbod):
            for cc in range(3):
                temp = max(temp, abs(F._v(b[q][cc][6])))
        temp = temp / (72.0 * math.pow(abs(t), 7.0))
        tdid = t
        if temp == 0.0:
            t = tdid * 1.4
        else:
            t = math.copysign(math.pow(tol / temp, 1.0 / 9.0), tdid)
        if st.dtflag == 1 and abs(t / tdid) < 1.0:
            t = t * .8
            continue
        if abs(t / tdid) > 1.4:
            t = tdid * 1.4
        break
    temp = tdid * tdid
    xc, vc = RA_XC, RA_VC
    for q in range(1, nbod):
        for cc in range(3):
            bk = b[q][cc]
            x1[q][cc] = (xc[7] * bk[6] + xc[6] * bk[5] + xc[5] * bk[4] + xc[4] * bk[3] + xc[3] * bk[2]
                         + xc[2] * bk[1] + xc[1] * bk[0] + xc[0] * a1[q][cc]) * temp + v1[q][cc] * tdid + x1[q][cc]
            v1[q][cc] = (vc[6] * bk[6] + vc[5] * bk[5] + vc[4] * bk[4] + vc[3] * bk[3] + vc[2] * bk[2]
                         + vc[1] * bk[1] + vc[0] * bk[0] + a1[q][cc]) * tdid + v1[q][cc]
    qq = t / tdid
    q2 = qq * qq; q3 = qq * q2; q4 = q2 * q2; q5 = q2 * q3; q6 = q3 * q3; q7 = q3 * q4
    for q in range(1, nbod):
        for cc in range(3):
            bk, ek = b[q][cc], e[q][cc]
            sd = [0.0] * 7 if st.dtflag == 1 else [bk[i] - ek[i] for i in range(7)]
            ek[0] = qq * (bk[6] * 7.0 + bk[5] * 6.0 + bk[4] * 5.0 + bk[3] * 4.0 + bk[2] * 3.0 + bk[1] * 2.0 + bk[0])
            ek[1] = q2 * (bk[6] * 21.0 + bk[5] * 15.0 + bk[4] * 10.0 + bk[3] * 6.0 + bk[2] * 3.0 + bk[1])
            ek[2] = q3 * (bk[6] * 35.0 + bk[5] * 20.0 + bk[4] * 10.0 + bk[3] * 4.0 + bk[2])
            ek[3] = q4 * (bk[6] * 35.0 + bk[5] * 15.0 + bk[4] * 5.0 + bk[3])
            ek[4] = q5 * (bk[6] * 21.0 + bk[5] * 6.0 + bk[4])
            ek[5] = q6 * (bk[6] * 7.0 + bk[5])
            ek[6] = q7 * bk[6]
            for i in range(7):
                bk[i] = ek[i] + sd[i]
    st.dtflag = 2
    return tdid, t, niter


def ra15_rows(sysd, nparticles):
    """The system's Big bodies (floats) and its first particles (F rows)."""
    nb = sysd["nbig"]
    sel = list(range(nb)) + list(range(nb, nb + nparticles))
    m = [float(sysd["m"][i]) for i in sel]
    x = [[float(q) for q in sysd["x"][i]] if i < nb else [F(q) for q in sysd["x"][i]] for i in sel]
    v = [[float(q) for q in sysd["v"][i]] if i < nb else [F(q) for q in sysd["v"][i]] for i in sel]
    return m, nb, x, v, sel


def ra15_measure(sysd, nparticles=8, nseq=4, tol=1e-12):
    """Counted operations per particle for each of nseq sequences from the
    system's trial step (the first is mdt_ra15's reset sequence)."""
    m, nb, x, v, _ = ra15_rows(sysd, nparticles)
    st = Ra15State(len(m))
    t = float(sysd["h"])
    out = []
    for _ in range(nseq):
        OPS[0] = 0
        tdid, t, niter = ra15_sequence(m, nb, x, v, t, tol, st)
        out.append((niter, OPS[0] / nparticles, tdid))
    return out


def ra15_verify(sysd, nparticles=8, nseq=3, tol=1e-12):
    """The counted RA15 restatement reproduces the oracle's RA15 with 1PN
    (orc_ra15_run) bit for bit over nseq sequences: states and sizes."""
    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    from oracle.bindings import Oracle
    m, nb, x, v, sel = ra15_rows(sysd, nparticles)
    sub = {k: np.asarray(sysd[k])[sel].copy() for k in ("m", "x", "v")}
    sub.update(nbig=nb, h=float(sysd["h"]))
    xo, vo, hdid, hn, _ = Oracle().ra15_run(sub, nseq, tol=tol, pn=1)
    st = Ra15State(len(m))
    t = float(sysd["h"])
    sizes = []
    for _ in range(nseq):
        tdid, t, _ = ra15_sequence(m, nb, x, v, t, tol, st)
        sizes.append(tdid)
    got_x = np.array([[F._v(q) for q in row] for row in x])
    got_v = np.array([[F._v(q) for q in row] for row in v])
    return (np.array_equal(got_x[1:], xo[1:]) and np.array_equal(got_v[1:], vo[1:])
            and np.array_equal(np.array(sizes), hdid) and t == hn)


if __name__ == "__main__":
    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    from mercuda_amd import systems
    s = systems.config_large(seed=2, nparticles=20000)
    print("counted restatement matches the oracle drift:", verify(s))
    r = measure(s)
    for k, (avg, n) in r.items():
        print(f"{k:6s} {avg:8.1f} ops/particle-step over {n} particle-steps")
    print("counted RA15 restatement matches the oracle RA15 with 1PN:", ra15_verify(s))
    for niter, ops, tdid in ra15_measure(s):
        print(f"RA15 sequence, {niter} iterations, {tdid:.4f} d: {ops:.1f} ops per particle")
