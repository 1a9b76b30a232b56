// Universal-variable Kepler drift for gfx950, FP64.
//
// Device restatement of the SWIFT drift chain that MERCURY6 uses
// (drift_one -> drift_dan -> drift_kepmd | drift_kepu{guess,new,lag,p3solve,
// stumpff,fchk}, mco_sine; mercury6_2.for:7602-8352, 2940-2983).  Every
// expression keeps the Fortran's left-to-right association and the library is
// compiled with -ffp-contract=off, so with IEEE division and square root the
// results match the CPU path bit for bit except where it calls libm (cos in
// the elliptic guess; fmod is exact on both sides).
//
// The chain is split for the particle kernels:
//   drift_fast()  -- the closed kepmd branch only; returns false (state
//                    untouched) when drift_dan would take any other branch or
//                    fail, so the caller can hand the particle to the general
//                    kernel;
//   drift_one()   -- the whole chain, including drift_one's ten-substep retry.
#pragma once
#include <hip/hip_runtime.h>

#include "rn_ops.hpp"

namespace mhip {

constexpr double kDanbyB = 1.0e-13;          // swift.inc:23
constexpr int kNlag2 = 400;                  // swift.inc:27
constexpr double kTwoPi = 2.0 * 3.141592653589793;   // swift.inc:35

__device__ __forceinline__ double fsign(double a, double b) { return copysign(fabs(a), b); }

// drift_kepu_stumpff (mercury6_2.for:8213-8277); x is updated in place.
__device__ inline void stumpff(double& x, double& c0, double& c1, double& c2, double& c3) {
    int n = 0;
    while (fabs(x) >= 0.1) { n = n + 1; x = x * .25; }
    double x2 = x * x, x3 = x * x2, x4 = x2 * x2, x5 = x2 * x3, x6 = x3 * x3;
    c2 = 1.147074559772972e-11 * x6 - 2.087675698786810e-9 * x5
       + 2.755731922398589e-7 * x4 - 2.480158730158730e-5 * x3
       + 1.388888888888889e-3 * x2 - 4.166666666666667e-2 * x + .5;
    c3 = 7.647163731819816e-13 * x6 - 1.605904383682161e-10 * x5
       + 2.505210838544172e-8 * x4 - 2.755731922398589e-6 * x3
       + 1.984126984126984e-4 * x2 - 8.333333333333333e-3 * x
       + 1.666666666666667e-1;
    c1 = 1.0 - x * c3;
    c0 = 1.0 - x * c2;
    for (int i = n; i >= 1; i--) {
        c3 = (c2 + c0 * c3) * .25;
        c2 = c1 * c1 * .5;
        c1 = c0 * c1;
        c0 = 2.0 * c0 * c0 - 1.0;
        x = x * 4.0;
    }
}

// drift_kepu_fchk (mercury6_2.for:7876-7907)
__device__ inline double kepu_fchk(double dt, double r0, double mu, double alpha, double u, double s) {
    double c0, c1, c2, c3;
    double x = s * s * alpha;
    stumpff(x, c0, c1, c2, c3);
    c1 = c1 * s;
    c2 = c2 * s * s;
    c3 = c3 * s * s * s;
    return r0 * c1 + u * c2 + mu * c3 - dt;
}

// drift_kepu_p3solve (mercury6_2.for:8140-8203).  The Fortran's a2**3d0,
// q**3d0 and r**2d0 are repeated products in the flang build of the reference.
__device__ inline int kepu_p3solve(double dt, double r0, double mu, double alpha, double u, double& s) {
    double denom = (mu - alpha * r0) / 6.0;
    double a2 = 0.5 * u / denom;
    double a1 = r0 / denom;
    double a0 = -dt / denom;
    double q = (a1 - a2 * a2 / 3.0) / 3.0;
    double r = (a1 * a2 - 3.0 * a0) / 6.0 - (a2 * a2 * a2) / 27.0;
    double sq2 = (q * q * q) + (r * r);
    if (sq2 >= 0.0) {
        double sq = sqrt(sq2), p1, p2;
        const double third = 1.0 / 3.0;
        if ((r + sq) <= 0.0) p1 = -pow(-(r + sq), third);
        else p1 = pow(r + sq, third);
        if ((r - sq) <= 0.0) p2 = -pow(-(r - sq), third);
        else p2 = pow(r - sq, third);
        s = p1 + p2 - a2 / 3.0;
        return 0;
    }
    s = 0.0;
    return 1;
}

// drift_kepu_guess (mercury6_2.for:7909-7982) with mco_sine inlined
// (mercury6_2.for:2940-2983): the reduced angle feeds the guess.
__device__ inline void kepu_guess(double dt, double r0, double mu, double alpha, double u, double& s) {
    if (alpha > 0.0) {
        if (dt / r0 <= 0.4) {
            s = dt / r0 - (dt * dt * u) / (2.0 * r0 * r0 * r0);
            return;
        }
        double a = mu / alpha;
        double en = sqrt(mu / (a * a * a));
        double ec = 1.0 - r0 / a;
        double es = u / (en * a * a);
        double e = sqrt(ec * ec + es * es);
        double y = en * dt - es;
        const double pi = 3.141592653589793, twopi = 2.0 * pi;
        if (y > 0.0) y = fmod(y, twopi);
        else y = fmod(y, twopi) + twopi;
        double cy = cos(y);
        double sy = (y > pi) ? -sqrt(1.0 - cy * cy) : sqrt(1.0 - cy * cy);
        double sigma = fsign(1.0, es * cy + ec * sy);
        double x = y + sigma * .85 * e;
        s = x / sqrt(alpha);
    } else {
        if (kepu_p3solve(dt, r0, mu, alpha, u, s) != 0) s = dt / r0;
    }
}

// drift_kepu_new (mercury6_2.for:8073-8138)
__device__ inline int kepu_new(double& s, double dt, double r0, double mu, double alpha, double u,
                               double& fp, double& c1, double& c2, double& c3) {
    double c0;
    for (int nc = 0; nc <= 6; nc++) {
        double s2 = s * s;
        double x = s2 * alpha;
        stumpff(x, c0, c1, c2, c3);
        c1 = c1 * s;
        c2 = c2 * s2;
        c3 = c3 * s * s2;
        double f = r0 * c1 + u * c2 + mu * c3 - dt;
        fp = r0 * c0 + u * c1 + mu * c2;
        double fpp = (mu - r0 * alpha) * c1 + u * c0;
        double fppp = (mu - r0 * alpha) * c0 - u * alpha * c1;
        double ds = -f / fp;
        ds = -f / (fp + .5 * ds * fpp);
        ds = -f / (fp + .5 * ds * fpp + ds * ds * fppp * .1666666666666667);
        s = s + ds;
        double fdt = f / dt;
        if (fdt * fdt < kDanbyB * kDanbyB) return 0;
    }
    return 1;
}

// drift_kepu_lag (mercury6_2.for:7984-8071)
__device__ inline int kepu_lag(double& s, double dt, double r0, double mu, double alpha, double u,
                               double& fp, double& c1, double& c2, double& c3) {
    const double ln = 5.0;
    double c0;
    for (int nc = 0; nc <= kNlag2; nc++) {
        double x = s * s * alpha;
        stumpff(x, c0, c1, c2, c3);
        c1 = c1 * s;
        c2 = c2 * s * s;
        c3 = c3 * s * s * s;
        double f = r0 * c1 + u * c2 + mu * c3 - dt;
        fp = r0 * c0 + u * c1 + mu * c2;
        double fpp = (-40.0 * alpha + mu) * c1 + u * c0;
        double ds = -ln * f / (fp + fsign(1.0, fp) * sqrt(fabs((ln - 1.0) * (ln - 1.0) * fp * fp
                                                             - (ln - 1.0) * ln * f * fpp)));
        s = s + ds;
        double fdt = f / dt;
        if (fdt * fdt < kDanbyB * kDanbyB) return 0;
    }
    return 2;
}

// drift_kepu (mercury6_2.for:7830-7874)
__device__ inline int kepu(double dt, double r0, double mu, double alpha, double u,
                           double& fp, double& c1, double& c2, double& c3) {
    double s;
    kepu_guess(dt, r0, mu, alpha, u, s);
    double st = s;
    int iflg = kepu_new(s, dt, r0, mu, alpha, u, fp, c1, c2, c3);
    if (iflg != 0) {
        double fo = kepu_fchk(dt, r0, mu, alpha, u, st);
        double fn = kepu_fchk(dt, r0, mu, alpha, u, s);
        if (fabs(fo) < fabs(fn)) s = st;
        iflg = kepu_lag(s, dt, r0, mu, alpha, u, fp, c1, c2, c3);
    }
    return iflg;
}

// drift_kepmd (mercury6_2.for:7754-7808).  Ops: IeeeOps, or RangeOps where
// the caller has bounded the orbit (drift_fast).
template <class Ops = IeeeOps>
__device__ __forceinline__ void kepmd(double dm, double es, double ec, double& x, double& s, double& c) {
    const double A0 = 39916800.0, A1 = 6652800.0, A2 = 332640.0, A3 = 7920.0, A4 = 110.0;
    constexpr double rA0 = 1.0 / 39916800.0, r3 = 1.0 / 3.0;
    double fac1 = Ops::rcp(1.0 - ec);
    double q = fac1 * dm;
    double fac2 = es * es * fac1 - Ops::div_c(ec, 3.0, r3);
    x = q * (1.0 - 0.5 * fac1 * q * (es - q * fac2));
    double y = x * x;
    s = Ops::div_c(x * (A0 - y * (A1 - y * (A2 - y * (A3 - y * (A4 - y))))), A0, rA0);
    c = Ops::sqrt(1.0 - s * s);
    double f = x - ec * s + es * (1.0 - c) - dm;
    double fp = 1. - ec * c + es * s;
    double fpp = ec * s + es * c;
    double fppp = ec * c - es * s;
    double dx = Ops::div(-f, fp);
    dx = Ops::div(-f, fp + 0.5 * dx * fpp);
    dx = Ops::div(-f, fp + 0.5 * dx * fpp + 0.16666666666666666 * dx * dx * fppp);
    x = x + dx;
    y = x * x;
    s = Ops::div_c(x * (A0 - y * (A1 - y * (A2 - y * (A3 - y * (A4 - y))))), A0, rA0);
    c = Ops::sqrt(1.0 - s * s);
}

struct Orbit { double x, y, z, vx, vy, vz; };

// The kepmd branch of drift_dan (mercury6_2.for:7628-7690).  Returns true and
// advances the state when drift_dan(dt) takes that branch and passes its
// fchk test; otherwise returns false with the state untouched.
//
// Divides and square roots use RangeOps (rn_ops.hpp).  Their operands stay in
// its exact range when r0^2 lies in [2^-200, 2^200] and mu, |dt| in
// [2^-100, 2^100] (the caller's gate): a lies in [r0/2, 2^53 r0] because a
// positive alpha = 2mu/r0 - v^2 is at least 2^-54 * 2mu/r0, so every
// quotient below stays within 2^+-600; inside the kepmd branch |ec|, |es| <=
// 0.6, |dm| <= 0.4.  The one unbounded numerator, u in es = u / (en a^2),
// keeps the IEEE divide.  A state outside the gate returns false, so the
// general kernel's IEEE chain takes it.
// rr: x.x of the state, x*x + y*y + z*z in that order (the caller's copy).
// The quotients by r0 (2mu/r0, a/r0) and by en (dm/en, (s - x)/en) share
// their divisor's reciprocal refinement (rn_ops.hpp).
__device__ __forceinline__ bool drift_fast(double mu, Orbit& o, double dt, double rr) {
    using Ops = RangeOps;
    constexpr double rTwoPi = 1.0 / kTwoPi;
    if (!(rr >= 0x1p-200 && rr <= 0x1p200)) return false;
    double r0 = Ops::sqrt(rr);
    const double rr0 = rcp_refine(r0);
    double v0s = o.vx * o.vx + o.vy * o.vy + o.vz * o.vz;
    double u = o.x * o.vx + o.y * o.vy + o.z * o.vz;
    double alpha = div_rn_r(2.0 * mu, r0, rr0) - v0s;
    if (!(alpha > 0.0)) return false;
    double a = Ops::div(mu, alpha);
    double asq = a * a;
    double en = Ops::sqrt(Ops::div(mu, a * asq));
    const double ren = rcp_refine(en);
    double ec = 1.0 - Ops::div(r0, a);
    double es = u / (en * asq);
    double esq = ec * ec + es * es;
    double dm = dt * en - (double)((int)Ops::div_c(dt * en, kTwoPi, rTwoPi)) * kTwoPi;
    double dtr = div_rn_r(dm, en, ren);
    if ((dm * dm > 0.16) || (esq > 0.36) || !(esq * dm * dm < 0.0016)) return false;
    double xkep, s, c;
    kepmd<Ops>(dm, es, ec, xkep, s, c);
    double fchk = (xkep - ec * s + es * (1.0 - c) - dm);
    if (fchk * fchk > kDanbyB) return false;
    double fp = 1.0 - ec * c + es * s;
    double f = div_rn_r(a, r0, rr0) * (c - 1.0) + 1.0;
    double g = dtr + div_rn_r(s - xkep, en, ren);
    double fdot = -Ops::div(a, r0 * fp) * en * s;
    double gdot = Ops::div(c - 1., fp) + 1.;
    double nx = o.x * f + o.vx * g, ny = o.y * f + o.vy * g, nz = o.z * f + o.vz * g;
    double nvx = o.x * fdot + o.vx * gdot, nvy = o.y * fdot + o.vy * gdot, nvz = o.z * fdot + o.vz * gdot;
    o.x = nx; o.y = ny; o.z = nz; o.vx = nvx; o.vy = nvy; o.vz = nvz;
    return true;
}

// drift_dan (mercury6_2.for:7602-7752): returns iflg.  *kep (when given)
// is set when the drift took the kepmd branch and passed its fchk test.
__device__ inline int drift_dan(double mu, Orbit& o, double dt0, bool* kep = nullptr) {
    double dt = dt0;
    double r0 = sqrt(o.x * o.x + o.y * o.y + o.z * o.z);
    double v0s = o.vx * o.vx + o.vy * o.vy + o.vz * o.vz;
    double u = o.x * o.vx + o.y * o.vy + o.z * o.vz;
    double alpha = 2.0 * mu / r0 - v0s;
    double f, g, fdot, gdot, fp, c1, c2, c3;
    if (alpha > 0.0) {
        double a = mu / alpha;
        double asq = a * a;
        double en = sqrt(mu / (a * asq));
        double ec = 1.0 - r0 / a;
        double es = u / (en * asq);
        double esq = ec * ec + es * es;
        double dm = dt * en - (double)((int)(dt * en / kTwoPi)) * kTwoPi;
        dt = dm / en;
        if (!((dm * dm > 0.16) || (esq > 0.36)) && (esq * dm * dm < 0.0016)) {
            double xkep, s, c;
            kepmd(dm, es, ec, xkep, s, c);
            double fchk = (xkep - ec * s + es * (1.0 - c) - dm);
            if (fchk * fchk > kDanbyB) return 1;
            fp = 1.0 - ec * c + es * s;
            f = (a / r0) * (c - 1.0) + 1.0;
            g = dt + (s - xkep) / en;
            fdot = -(a / (r0 * fp)) * en * s;
            gdot = (c - 1.) / fp + 1.;
            double nx = o.x * f + o.vx * g, ny = o.y * f + o.vy * g, nz = o.z * f + o.vz * g;
            double nvx = o.x * fdot + o.vx * gdot, nvy = o.y * fdot + o.vy * gdot,
                   nvz = o.z * fdot + o.vz * gdot;
            o.x = nx; o.y = ny; o.z = nz; o.vx = nvx; o.vy = nvy; o.vz = nvz;
            if (kep) *kep = true;
            return 0;
        }
    }
    int iflg = kepu(dt, r0, mu, alpha, u, fp, c1, c2, c3);
    if (iflg == 0) {
        f = 1.0 - (mu / r0) * c2;
        g = dt - mu * c3;
        fdot = -(mu / (fp * r0)) * c1;
        gdot = 1.0 - (mu / fp) * c2;
        double nx = o.x * f + o.vx * g, ny = o.y * f + o.vy * g, nz = o.z * f + o.vz * g;
        double nvx = o.x * fdot + o.vx * gdot, nvy = o.y * fdot + o.vy * gdot, nvz = o.z * fdot + o.vz * gdot;
        o.x = nx; o.y = ny; o.z = nz; o.vx = nvx; o.vy = nvy; o.vz = nvz;
    }
    return iflg;
}

// drift_one's retry of a failed drift as ten drifts of dt/10, stopping at the
// first failure with partial progress kept (not inlined: rarely taken).
__device__ __noinline__ int drift_retry(double mu, Orbit& o, double dt) {
    int iflg = 0;
    for (int i = 1; i <= 10; i++) {
        double dttmp = dt / 10.0;
        iflg = drift_dan(mu, o, dttmp);
        if (iflg != 0) return iflg;
    }
    return iflg;
}

// drift_one (mercury6_2.for:8279-8352).  *kep (when given) reports whether the
// first drift_dan took the kepmd branch (then no retry follows): the drift
// drift_fast would have made.
__device__ inline int drift_one(double mu, Orbit& o, double dt, bool* kep = nullptr) {
    if (kep) *kep = false;
    int iflg = drift_dan(mu, o, dt, kep);
    if (iflg != 0) iflg = drift_retry(mu, o, dt);
    return iflg;
}

}  // namespace mhip
