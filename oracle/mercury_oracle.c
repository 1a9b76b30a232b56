/*
 * mercury_oracle.c -- CPU restatement of MERCURY6's massless-body hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or shipped).
 *
 * Each function follows one routine of mercury6_2.for line by line, keeping the
 * Fortran's left-to-right association and operation order, so that
 *     gcc -O2 -ffp-contract=off
 * reproduces the flang-built reference bit for bit except where both call
 * libm.  File:line citations are to the reference repository.
 */
#include "mercury_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Gaussian gravitational constant squared (mercury.inc:37) */
#define K2 2.959122082855911e-4

/* the non-gravitational model (orc_set_ngf) */
static int g_ngflag = 0;
static const double *g_ngf = NULL;
static double *g_angf = NULL;

/* swift.inc:22-31, mercury.inc:37-44 */
#define DANBYB 1.0e-13
#define NLAG2 400
static const double SW_TWOPI = 2.0 * 3.141592653589793;
static const double HUGEV = 9.9e29;

static double fsign(double a, double b) { return copysign(fabs(a), b); }

/* ------------------------------------------------------------------------ */
/* mco_sine  mercury6_2.for:2940-2983.  x is modified in place, as in the
 * Fortran (drift_kepu_guess relies on the reduced value). */
static void mco_sine_inplace(double *x, double *sx, double *cx)
{
    const double pi = 3.141592653589793;
    const double twopi = 2.0 * pi;
    if (*x > 0.0)
        *x = fmod(*x, twopi);
    else
        *x = fmod(*x, twopi) + twopi;
    *cx = cos(*x);
    if (*x > pi)
        *sx = -sqrt(1.0 - *cx * *cx);
    else
        *sx = sqrt(1.0 - *cx * *cx);
}

void orc_mco_sine(double x, double *sx, double *cx) { mco_sine_inplace(&x, sx, cx); }

/* drift_kepu_stumpff  mercury6_2.for:8213-8277 */
static void stumpff(double *xp, double *c0, double *c1, double *c2, double *c3)
{
    double x = *xp;
    int n = 0;
    const double xm = 0.1;
    while (fabs(x) >= xm) {
        n = n + 1;
        x = x * .25;
    }
    double x2 = x * x;
    double x3 = x * x2;
    double x4 = x2 * x2;
    double x5 = x2 * x3;
    double x6 = x3 * x3;
    *c2 = 1.147074559772972e-11 * x6 - 2.087675698786810e-9 * x5
        + 2.755731922398589e-7 * x4 - 2.480158730158730e-5 * x3
        + 1.388888888888889e-3 * x2 - 4.166666666666667e-2 * x + .5;
    *c3 = 7.647163731819816e-13 * x6 - 1.605904383682161e-10 * x5
        + 2.505210838544172e-8 * x4 - 2.755731922398589e-6 * x3
        + 1.984126984126984e-4 * x2 - 8.333333333333333e-3 * x
        + 1.666666666666667e-1;
    *c1 = 1.0 - x * *c3;
    *c0 = 1.0 - x * *c2;
    if (n != 0) {
        for (int i = n; i >= 1; i--) {
            *c3 = (*c2 + *c0 * *c3) * .25;
            *c2 = *c1 * *c1 * .5;
            *c1 = *c0 * *c1;
            *c0 = 2.0 * *c0 * *c0 - 1.0;
            x = x * 4.0;
        }
    }
    *xp = x;
}

void orc_drift_kepu_stumpff(double x, double *c) { stumpff(&x, &c[0], &c[1], &c[2], &c[3]); }

/* drift_kepu_fchk  mercury6_2.for:7876-7907 */
static double kepu_fchk(double dt, double r0, double mu, double alpha, double u, double s)
{
    double c0, c1, c2, c3;
    double x = s * s * alpha;
    stumpff(&x, &c0, &c1, &c2, &c3);
    c1 = c1 * s;
    c2 = c2 * s * s;
    c3 = c3 * s * s * s;
    return r0 * c1 + u * c2 + mu * c3 - dt;
}

/* drift_kepu_p3solve  mercury6_2.for:8140-8203 */
static int kepu_p3solve(double dt, double r0, double mu, double alpha, double u, double *s)
{
    double denom = (mu - alpha * r0) / 6.0;
    double a2 = 0.5 * u / denom;
    double a1 = r0 / denom;
    double a0 = -dt / denom;
    double q = (a1 - a2 * a2 / 3.0) / 3.0;
    /* a2**3d0, q**3d0, r**2d0: the flang build of the reference evaluates a
     * real power with an integral constant exponent as repeated products */
    double r = (a1 * a2 - 3.0 * a0) / 6.0 - (a2 * a2 * a2) / 27.0;
    double sq2 = (q * q * q) + (r * r);
    if (sq2 >= 0.0) {
        double sq = sqrt(sq2), p1, p2;
        if ((r + sq) <= 0.0)
            p1 = -pow(-(r + sq), 1.0 / 3.0);
        else
            p1 = pow(r + sq, 1.0 / 3.0);
        if ((r - sq) <= 0.0)
            p2 = -pow(-(r - sq), 1.0 / 3.0);
        else
            p2 = pow(r - sq, 1.0 / 3.0);
        *s = p1 + p2 - a2 / 3.0;
        return 0;
    }
    *s = 0.0;
    return 1;
}

/* drift_kepu_guess  mercury6_2.for:7909-7982 */
static void kepu_guess(double dt, double r0, double mu, double alpha, double u, double *s)
{
    if (alpha > 0.0) {
        if (dt / r0 <= 0.4) {
            *s = dt / r0 - (dt * dt * u) / (2.0 * r0 * r0 * r0);
            return;
        }
        double a = mu / alpha;
        double en = sqrt(mu / (a * a * a));
        double ec = 1.0 - r0 / a;
        double es = u / (en * a * a);
        double e = sqrt(ec * ec + es * es);
        double y = en * dt - es;
        double sy, cy;
        mco_sine_inplace(&y, &sy, &cy);
        double sigma = fsign(1.0, es * cy + ec * sy);
        double x = y + sigma * .85 * e;
        *s = x / sqrt(alpha);
    } else {
        if (kepu_p3solve(dt, r0, mu, alpha, u, s) != 0)
            *s = dt / r0;
    }
}

/* drift_kepu_new  mercury6_2.for:8073-8138 */
static int kepu_new(double *s, double dt, double r0, double mu, double alpha, double u,
                    double *fp, double *c1, double *c2, double *c3)
{
    double c0;
    for (int nc = 0; nc <= 6; nc++) {
        double s2 = *s * *s;
        double x = s2 * alpha;
        stumpff(&x, &c0, c1, c2, c3);
        *c1 = *c1 * *s;
        *c2 = *c2 * s2;
        *c3 = *c3 * *s * s2;
        double f = r0 * *c1 + u * *c2 + mu * *c3 - dt;
        *fp = r0 * c0 + u * *c1 + mu * *c2;
        double fpp = (mu - r0 * alpha) * *c1 + u * c0;
        double fppp = (mu - r0 * alpha) * c0 - u * alpha * *c1;
        double ds = -f / *fp;
        ds = -f / (*fp + .5 * ds * fpp);
        ds = -f / (*fp + .5 * ds * fpp + ds * ds * fppp * .1666666666666667);
        *s = *s + ds;
        double fdt = f / dt;
        if (fdt * fdt < DANBYB * DANBYB)
            return 0;
    }
    return 1;
}

/* drift_kepu_lag  mercury6_2.for:7984-8071 */
static int kepu_lag(double *s, double dt, double r0, double mu, double alpha, double u,
                    double *fp, double *c1, double *c2, double *c3)
{
    const double ln = 5.0;
    double c0;
    for (int nc = 0; nc <= NLAG2; nc++) {
        double x = *s * *s * alpha;
        stumpff(&x, &c0, c1, c2, c3);
        *c1 = *c1 * *s;
        *c2 = *c2 * *s * *s;
        *c3 = *c3 * *s * *s * *s;
        double f = r0 * *c1 + u * *c2 + mu * *c3 - dt;
        *fp = r0 * c0 + u * *c1 + mu * *c2;
        double fpp = (-40.0 * alpha + mu) * *c1 + u * c0;
        double ds = -ln * f / (*fp + fsign(1.0, *fp) * sqrt(fabs((ln - 1.0) * (ln - 1.0) * *fp * *fp
                                                            - (ln - 1.0) * ln * f * fpp)));
        *s = *s + ds;
        double fdt = f / dt;
        if (fdt * fdt < DANBYB * DANBYB)
            return 0;
    }
    return 2;
}

/* drift_kepu  mercury6_2.for:7830-7874 */
static int kepu(double dt, double r0, double mu, double alpha, double u,
                double *fp, double *c1, double *c2, double *c3)
{
    double s;
    kepu_guess(dt, r0, mu, alpha, u, &s);
    double st = s;
    int iflg = kepu_new(&s, dt, r0, mu, alpha, u, fp, c1, c2, c3);
    if (iflg != 0) {
        double fo = kepu_fchk(dt, r0, mu, alpha, u, st);
        double fn = kepu_fchk(dt, r0, mu, alpha, u, s);
        if (fabs(fo) < fabs(fn))
            s = st;
        iflg = kepu_lag(&s, dt, r0, mu, alpha, u, fp, c1, c2, c3);
    }
    return iflg;
}

/* drift_kepmd  mercury6_2.for:7754-7808 */
static void kepmd(double dm, double es, double ec, double *x, double *s, double *c)
{
    const double A0 = 39916800.0, A1 = 6652800.0, A2 = 332640.0, A3 = 7920.0, A4 = 110.0;
    double fac1 = 1.0 / (1.0 - ec);
    double q = fac1 * dm;
    double fac2 = es * es * fac1 - ec / 3.0;
    *x = q * (1.0 - 0.5 * fac1 * q * (es - q * fac2));
    double y = *x * *x;
    *s = *x * (A0 - y * (A1 - y * (A2 - y * (A3 - y * (A4 - y))))) / A0;
    *c = sqrt(1.0 - *s * *s);
    double f = *x - ec * *s + es * (1.0 - *c) - dm;
    double fp = 1. - ec * *c + es * *s;
    double fpp = ec * *s + es * *c;
    double fppp = ec * *c - es * *s;
    double dx = -f / fp;
    dx = -f / (fp + 0.5 * dx * fpp);
    dx = -f / (fp + 0.5 * dx * fpp + 0.16666666666666666 * dx * dx * fppp);
    *x = *x + dx;
    y = *x * *x;
    *s = *x * (A0 - y * (A1 - y * (A2 - y * (A3 - y * (A4 - y))))) / A0;
    *c = sqrt(1.0 - *s * *s);
}

/* drift_dan  mercury6_2.for:7602-7752.  branch_out (optional) reports the
 * path taken: 0 kepmd, 1 kepu. */
static int drift_dan_impl(double mu, double *xv, double dt0, int *branch_out)
{
    double x0 = xv[0], y0 = xv[1], z0 = xv[2], vx0 = xv[3], vy0 = xv[4], vz0 = xv[5];
    double dt = dt0;
    double f, g, fdot, gdot, c1, c2, c3, fp;
    double r0 = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
    double v0s = vx0 * vx0 + vy0 * vy0 + vz0 * vz0;
    double u = x0 * vx0 + y0 * vy0 + z0 * vz0;
    double alpha = 2.0 * mu / r0 - v0s;

    if (alpha > 0.0) {
        double a = mu / alpha;
        double asq = a * a;
        double en = sqrt(mu / (a * asq));
        double ec = 1.0 - r0 / a;
        double es = u / (en * asq);
        double esq = ec * ec + es * es;
        double dm = dt * en - (double)((int)(dt * en / SW_TWOPI)) * SW_TWOPI;
        dt = dm / en;
        if (!((dm * dm > 0.16) || (esq > 0.36)) && (esq * dm * dm < 0.0016)) {
            double xkep, s, c;
            if (branch_out) *branch_out = 0;
            kepmd(dm, es, ec, &xkep, &s, &c);
            double fchk = (xkep - ec * s + es * (1.0 - c) - dm);
            if (fchk * fchk > DANBYB)
                return 1;
            fp = 1.0 - ec * c + es * s;
            f = (a / r0) * (c - 1.0) + 1.0;
            g = dt + (s - xkep) / en;
            fdot = -(a / (r0 * fp)) * en * s;
            gdot = (c - 1.) / fp + 1.;
            xv[0] = x0 * f + vx0 * g;
            xv[1] = y0 * f + vy0 * g;
            xv[2] = z0 * f + vz0 * g;
            xv[3] = x0 * fdot + vx0 * gdot;
            xv[4] = y0 * fdot + vy0 * gdot;
            xv[5] = z0 * fdot + vz0 * gdot;
            return 0;
        }
    }
    if (branch_out) *branch_out = 1;
    int iflg = kepu(dt, r0, mu, alpha, u, &fp, &c1, &c2, &c3);
    if (iflg == 0) {
        f = 1.0 - (mu / r0) * c2;
        g = dt - mu * c3;
        fdot = -(mu / (fp * r0)) * c1;
        gdot = 1.0 - (mu / fp) * c2;
        xv[0] = x0 * f + vx0 * g;
        xv[1] = y0 * f + vy0 * g;
        xv[2] = z0 * f + vz0 * g;
        xv[3] = x0 * fdot + vx0 * gdot;
        xv[4] = y0 * fdot + vy0 * gdot;
        xv[5] = z0 * fdot + vz0 * gdot;
    }
    return iflg;
}

int orc_drift_dan(double mu, double *xv, double dt) { return drift_dan_impl(mu, xv, dt, NULL); }

/* drift_one  mercury6_2.for:8279-8352 */
int orc_drift_one(double mu, double *xv, double dt)
{
    int iflg = drift_dan_impl(mu, xv, dt, NULL);
    if (iflg != 0) {
        for (int i = 1; i <= 10; i++) {
            double dttmp = dt / 10.0;
            iflg = drift_dan_impl(mu, xv, dttmp, NULL);
            if (iflg != 0)
                return iflg;
        }
    }
    return iflg;
}

int orc_drift_branch(double mu, const double *xv, double dt)
{
    double w[6];
    memcpy(w, xv, sizeof w);
    int branch = 0;
    int iflg = drift_dan_impl(mu, w, dt, &branch);
    if (iflg != 0) return 3;
    if (branch == 0) return 0;
    /* distinguish Newton from Laguerre by re-running Newton alone */
    double x0 = xv[0], y0 = xv[1], z0 = xv[2], vx0 = xv[3], vy0 = xv[4], vz0 = xv[5];
    double r0 = sqrt(x0 * x0 + y0 * y0 + z0 * z0);
    double v0s = vx0 * vx0 + vy0 * vy0 + vz0 * vz0;
    double u = x0 * vx0 + y0 * vy0 + z0 * vz0;
    double alpha = 2.0 * mu / r0 - v0s;
    double dtl = dt;
    if (alpha > 0.0) {
        double a = mu / alpha, asq = a * a, en = sqrt(mu / (a * asq));
        double dm = dtl * en - (double)((int)(dtl * en / SW_TWOPI)) * SW_TWOPI;
        dtl = dm / en;
    }
    double s, fp, c1, c2, c3;
    kepu_guess(dtl, r0, mu, alpha, u, &s);
    return kepu_new(&s, dtl, r0, mu, alpha, u, &fp, &c1, &c2, &c3) == 0 ? 1 : 2;
}

/* ------------------------------------------------------------------------ */
/* mco_h2j  mercury6_2.for:2613-2687 (called with nbod = nbig by mdt_mvs) */
void orc_mco_h2j(int nbig, const double *m, const double *xh, const double *vh,
                 double *x, double *v)
{
    if (nbig < 2) return;
    /* Fortran body index b (1-based) -> array index b-1 */
    double mtot = m[1];
    for (int k = 0; k < 3; k++) { x[3 + k] = xh[3 + k]; v[3 + k] = vh[3 + k]; }
    double mx = m[1] * xh[3], my = m[1] * xh[4], mz = m[1] * xh[5];
    double mu = m[1] * vh[3], mv = m[1] * vh[4], mw = m[1] * vh[5];
    for (int j = 3; j <= nbig - 1; j++) {
        int b = j - 1;
        double temp = 1.0 / (mtot + m[0]);
        mtot = mtot + m[b];
        x[3 * b + 0] = xh[3 * b + 0] - temp * mx;
        x[3 * b + 1] = xh[3 * b + 1] - temp * my;
        x[3 * b + 2] = xh[3 * b + 2] - temp * mz;
        v[3 * b + 0] = vh[3 * b + 0] - temp * mu;
        v[3 * b + 1] = vh[3 * b + 1] - temp * mv;
        v[3 * b + 2] = vh[3 * b + 2] - temp * mw;
        mx = mx + m[b] * xh[3 * b + 0];
        my = my + m[b] * xh[3 * b + 1];
        mz = mz + m[b] * xh[3 * b + 2];
        mu = mu + m[b] * vh[3 * b + 0];
        mv = mv + m[b] * vh[3 * b + 1];
        mw = mw + m[b] * vh[3 * b + 2];
    }
    if (nbig > 2) {
        int b = nbig - 1;
        double temp = 1.0 / (mtot + m[0]);
        x[3 * b + 0] = xh[3 * b + 0] - temp * mx;
        x[3 * b + 1] = xh[3 * b + 1] - temp * my;
        x[3 * b + 2] = xh[3 * b + 2] - temp * mz;
        v[3 * b + 0] = vh[3 * b + 0] - temp * mu;
        v[3 * b + 1] = vh[3 * b + 1] - temp * mv;
        v[3 * b + 2] = vh[3 * b + 2] - temp * mw;
    }
}

/* mco_j2h  mercury6_2.for:2700-2770 */
void orc_mco_j2h(int nbig, const double *m, const double *x, const double *v,
                 double *xh, double *vh)
{
    if (nbig < 2) return;
    for (int k = 0; k < 3; k++) { xh[3 + k] = x[3 + k]; vh[3 + k] = v[3 + k]; }
    double mtot = m[1];
    double temp = m[1] / (mtot + m[0]);
    double mx = temp * x[3], my = temp * x[4], mz = temp * x[5];
    double mu = temp * v[3], mv = temp * v[4], mw = temp * v[5];
    for (int j = 3; j <= nbig - 1; j++) {
        int b = j - 1;
        xh[3 * b + 0] = x[3 * b + 0] + mx;
        xh[3 * b + 1] = x[3 * b + 1] + my;
        xh[3 * b + 2] = x[3 * b + 2] + mz;
        vh[3 * b + 0] = v[3 * b + 0] + mu;
        vh[3 * b + 1] = v[3 * b + 1] + mv;
        vh[3 * b + 2] = v[3 * b + 2] + mw;
        mtot = mtot + m[b];
        temp = m[b] / (mtot + m[0]);
        mx = mx + temp * x[3 * b + 0];
        my = my + temp * x[3 * b + 1];
        mz = mz + temp * x[3 * b + 2];
        mu = mu + temp * v[3 * b + 0];
        mv = mv + temp * v[3 * b + 1];
        mw = mw + temp * v[3 * b + 2];
    }
    if (nbig > 2) {
        int b = nbig - 1;
        xh[3 * b + 0] = x[3 * b + 0] + mx;
        xh[3 * b + 1] = x[3 * b + 1] + my;
        xh[3 * b + 2] = x[3 * b + 2] + mz;
        vh[3 * b + 0] = v[3 * b + 0] + mu;
        vh[3 * b + 1] = v[3 * b + 1] + mv;
        vh[3 * b + 2] = v[3 * b + 2] + mw;
    }
}

/* ------------------------------------------------------------------------ */
/* mfo_obl  mercury6_2.for:5176-5225 */
void orc_mfo_obl(const double *jcen, int nbod, const double *m, const double *x,
                 double *a, double *acen)
{
    acen[0] = acen[1] = acen[2] = 0.0;
    for (int i = 1; i < nbod; i++) {
        const double *xi = x + 3 * i;
        double r2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
        double r_1 = 1.0 / sqrt(r2);
        double r_2 = r_1 * r_1;
        double r_3 = r_2 * r_1;
        double jr2 = jcen[0] * r_2;
        double jr4 = jcen[1] * r_2 * r_2;
        double jr6 = jcen[2] * r_2 * r_2 * r_2;
        double u2 = xi[2] * xi[2] * r_2;
        double u4 = u2 * u2;
        double u6 = u4 * u2;
        double tmp1 = m[0] * r_3;
        double tmp2 = jr2 * (7.5 * u2 - 1.5)
                    + jr4 * (39.375 * u4 - 26.25 * u2 + 1.875)
                    + jr6 * (187.6875 * u6 - 216.5625 * u4 + 59.0625 * u2 - 2.1875);
        double tmp3 = jr2 * 3.0 + jr4 * (17.5 * u2 - 7.5)
                    + jr6 * (86.625 * u4 - 78.75 * u2 + 13.125);
        a[3 * i + 0] = xi[0] * tmp1 * tmp2;
        a[3 * i + 1] = xi[1] * tmp1 * tmp2;
        a[3 * i + 2] = xi[2] * tmp1 * (tmp2 - tmp3);
        double tmp4 = m[i] / m[0];
        acen[0] = acen[0] - tmp4 * a[3 * i + 0];
        acen[1] = acen[1] - tmp4 * a[3 * i + 1];
        acen[2] = acen[2] - tmp4 * a[3 * i + 2];
    }
}

/* mfo_mvs  mercury6_2.for:4929-5074 */
void orc_mfo_mvs(const double *jcen, int nbod, int nbig, const double *m,
                 const double *x, const double *xj, double *a)
{
    double a0[3] = {0.0, 0.0, 0.0}, a0tp[3] = {0.0, 0.0, 0.0};
    if (nbod < 2) return;
    double *a1 = calloc((size_t)3 * nbod, sizeof(double));
    double *a2 = calloc((size_t)3 * nbod, sizeof(double));
    double *a3 = calloc((size_t)3 * nbod, sizeof(double));
    double minside = 0.0;
    for (int kk = 3; kk <= nbig; kk++) {
        int k = kk - 1, k1 = kk - 2;
        minside = minside + m[k1];
        double r2 = x[3 * k] * x[3 * k] + x[3 * k + 1] * x[3 * k + 1] + x[3 * k + 2] * x[3 * k + 2];
        double rj2 = xj[3 * k] * xj[3 * k] + xj[3 * k + 1] * xj[3 * k + 1] + xj[3 * k + 2] * xj[3 * k + 2];
        double r = 1.0 / sqrt(r2);
        double rj = 1.0 / sqrt(rj2);
        double r3 = r * r * r;
        double rj3 = rj * rj * rj;
        double fac0 = m[k] * r3;
        double fac12 = m[0] * rj3;
        double fac2 = m[k] * fac12 / (minside + m[0]);
        double q = (r2 - rj2) * .5 / rj2;
        double q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q3 * q4;
        double fac1 = 402.1875 * q7 - 187.6875 * q6 + 86.625 * q5
                    - 39.375 * q4 + 17.5 * q3 - 7.5 * q2 + 3.0 * q - 1.0;
        for (int c = 0; c < 3; c++) {
            a0[c] = a0[c] - fac0 * x[3 * k + c];
            a1[3 * k + c] = fac12 * (xj[3 * k + c] + fac1 * x[3 * k + c]);
            a2[3 * k + c] = a2[3 * k1 + c] + fac2 * xj[3 * k + c];
        }
    }
    if (nbig >= 2) {
        double r2 = x[3] * x[3] + x[4] * x[4] + x[5] * x[5];
        double r = 1.0 / sqrt(r2);
        double r3 = r * r * r;
        double fac0 = m[1] * r3;
        for (int c = 0; c < 3; c++) a0tp[c] = a0[c] - fac0 * x[3 + c];
    }
    for (int i = 1; i < nbig; i++) {
        for (int j = i + 1; j < nbig; j++) {
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double s2 = dx * dx + dy * dy + dz * dz;
            double s_1 = 1.0 / sqrt(s2);
            double s_3 = s_1 * s_1 * s_1;
            double faci = m[i] * s_3;
            double facj = m[j] * s_3;
            a3[3 * j + 0] = a3[3 * j + 0] - faci * dx;
            a3[3 * j + 1] = a3[3 * j + 1] - faci * dy;
            a3[3 * j + 2] = a3[3 * j + 2] - faci * dz;
            a3[3 * i + 0] = a3[3 * i + 0] + facj * dx;
            a3[3 * i + 1] = a3[3 * i + 1] + facj * dy;
            a3[3 * i + 2] = a3[3 * i + 2] + facj * dz;
        }
        for (int j = nbig; j < nbod; j++) {
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double s2 = dx * dx + dy * dy + dz * dz;
            double s_1 = 1.0 / sqrt(s2);
            double s_3 = s_1 * s_1 * s_1;
            double faci = m[i] * s_3;
            a3[3 * j + 0] = a3[3 * j + 0] - faci * dx;
            a3[3 * j + 1] = a3[3 * j + 1] - faci * dy;
            a3[3 * j + 2] = a3[3 * j + 2] - faci * dz;
        }
    }
    for (int k = 1; k < nbig; k++)
        for (int c = 0; c < 3; c++)
            a[3 * k + c] = a0[c] + a1[3 * k + c] + a2[3 * k + c] + a3[3 * k + c];
    for (int k = nbig; k < nbod; k++)
        for (int c = 0; c < 3; c++)
            a[3 * k + c] = a0tp[c] + a3[3 * k + c];
    if (jcen[0] != 0 || jcen[1] != 0 || jcen[2] != 0) {
        double *aobl = calloc((size_t)3 * nbod, sizeof(double));
        double acen[3];
        orc_mfo_obl(jcen, nbod, m, x, aobl, acen);
        for (int k = 1; k < nbod; k++)
            for (int c = 0; c < 3; c++)
                a[3 * k + c] = a[3 * k + c] + (aobl[3 * k + c] - acen[c]);
        free(aobl);
    }
    free(a1);
    free(a2);
    free(a3);
}

/* ------------------------------------------------------------------------ */
/* mce_min  mercury6_2.for:1572-1624 */
void orc_mce_min(double d0, double d1, double d0t, double d1t, double h,
                 double *d2min, double *tmin)
{
    if (d0t * h > 0.0 || d1t * h < 0.0) {
        if (d0 <= d1) {
            *d2min = d0;
            *tmin = -h;
        } else {
            *d2min = d1;
            *tmin = 0.0;
        }
    } else {
        double temp = 6.0 * (d0 - d1);
        double a = temp + 3.0 * h * (d0t + d1t);
        double b = temp + 2.0 * h * (d0t + 2.0 * d1t);
        double c = h * d1t;
        temp = -.5 * (b + fsign(sqrt(fmax(b * b - 4.0 * a * c, 0.0)), b));
        double tau;
        if (temp == 0.0)
            tau = 0.0;
        else
            tau = c / temp;
        tau = fmin(tau, 0.0);
        tau = fmax(tau, -1.0);
        *tmin = tau * h;
        temp = 1.0 + tau;
        *d2min = tau * tau * ((3.0 + 2.0 * tau) * d0 + temp * h * d0t)
               + temp * temp * ((1.0 - 2.0 * tau) * d1 + tau * h * d1t);
        *d2min = fmax(*d2min, 0.0);
    }
}

/* mce_box  mercury6_2.for:997-1058 */
void orc_mce_box(int nbod, double h, const double *x0, const double *v0,
                 const double *x1, const double *v1, double *xmin, double *xmax,
                 double *ymin, double *ymax)
{
    for (int j = 1; j < nbod; j++) {
        xmin[j] = fmin(x0[3 * j], x1[3 * j]);
        xmax[j] = fmax(x0[3 * j], x1[3 * j]);
        ymin[j] = fmin(x0[3 * j + 1], x1[3 * j + 1]);
        ymax[j] = fmax(x0[3 * j + 1], x1[3 * j + 1]);
        if ((v0[3 * j] < 0.0 && v1[3 * j] > 0.0) || (v0[3 * j] > 0.0 && v1[3 * j] < 0.0)) {
            double temp = (v0[3 * j] * x1[3 * j] - v1[3 * j] * x0[3 * j]
                           - .5 * h * v0[3 * j] * v1[3 * j]) / (v0[3 * j] - v1[3 * j]);
            xmin[j] = fmin(xmin[j], temp);
            xmax[j] = fmax(xmax[j], temp);
        }
        if ((v0[3 * j + 1] < 0.0 && v1[3 * j + 1] > 0.0) || (v0[3 * j + 1] > 0.0 && v1[3 * j + 1] < 0.0)) {
            double temp = (v0[3 * j + 1] * x1[3 * j + 1] - v1[3 * j + 1] * x0[3 * j + 1]
                           - .5 * h * v0[3 * j + 1] * v1[3 * j + 1]) / (v0[3 * j + 1] - v1[3 * j + 1]);
            ymin[j] = fmin(ymin[j], temp);
            ymax[j] = fmax(ymax[j], temp);
        }
    }
}

/* mce_stat  mercury6_2.for:1765-1931 (all bodies active: stat = 0) */
void orc_mce_stat(double time, double h, double rcen, int nbod, int nbig,
                  const double *m, const double *x0, const double *v0,
                  const double *x1, const double *v1, const double *rce,
                  const double *rphys, orc_events *ev)
{
    (void)rcen;
    ev->nhit = 0;
    ev->thit1 = fsign(HUGEV, h);
    double hm1 = 1.0 / h;
    double *xmin = malloc(sizeof(double) * nbod), *xmax = malloc(sizeof(double) * nbod);
    double *ymin = malloc(sizeof(double) * nbod), *ymax = malloc(sizeof(double) * nbod);
    orc_mce_box(nbod, h, x0, v0, x1, v1, xmin, xmax, ymin, ymax);
    for (int j = 1; j < nbod; j++) {
        double temp = rce[j] * 1.2;
        xmin[j] = xmin[j] - temp;
        xmax[j] = xmax[j] + temp;
        ymin[j] = ymin[j] - temp;
        ymax[j] = ymax[j] + temp;
    }
    for (int i = 1; i < nbig; i++) {
        for (int j = i + 1; j < nbod; j++) {
            if (!(xmax[i] >= xmin[j] && xmax[j] >= xmin[i] && ymax[i] >= ymin[j] && ymax[j] >= ymin[i]))
                continue;
            double dx0 = x0[3 * i] - x0[3 * j];
            double dy0 = x0[3 * i + 1] - x0[3 * j + 1];
            double dz0 = x0[3 * i + 2] - x0[3 * j + 2];
            double du0 = v0[3 * i] - v0[3 * j];
            double dv0 = v0[3 * i + 1] - v0[3 * j + 1];
            double dw0 = v0[3 * i + 2] - v0[3 * j + 2];
            double d0t = (dx0 * du0 + dy0 * dv0 + dz0 * dw0) * 2.0;
            double dx1 = x1[3 * i] - x1[3 * j];
            double dy1 = x1[3 * i + 1] - x1[3 * j + 1];
            double dz1 = x1[3 * i + 2] - x1[3 * j + 2];
            double du1 = v1[3 * i] - v1[3 * j];
            double dv1 = v1[3 * i + 1] - v1[3 * j + 1];
            double dw1 = v1[3 * i + 2] - v1[3 * j + 2];
            double d1t = (dx1 * du1 + dy1 * dv1 + dz1 * dw1) * 2.0;
            double d0 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0;
            double d1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
            double d2min, tmin;
            orc_mce_min(d0, d1, d0t, d1t, h, &d2min, &tmin);
            double d2ce = fmax(rce[i], rce[j]);
            double d2hit = rphys[i] + rphys[j];
            d2ce = d2ce * d2ce;
            d2hit = d2hit * d2hit;
            double d2near = d2hit * 4.0;
            if ((d2min <= d2ce && d0t * h <= 0.0 && d1t * h >= 0.0) || (d2min <= d2hit)) {
                if (ev->nclo >= ev->cap) {
                    ev->overflow = 1;
                } else {
                    int k = ev->nclo++;
                    ev->tclo[k] = tmin + time;
                    ev->dclo[k] = sqrt(fmax(0.0, d2min));
                    int ic = i, jc = j;
                    if (m[j] > m[i] && j < nbig) { ic = j; jc = i; }
                    ev->iclo[k] = ic + 1;
                    ev->jclo[k] = jc + 1;
                    double tmp0 = -tmin * hm1;
                    double tmp1 = 1.0 + tmin * hm1;
                    for (int c = 0; c < 3; c++) {
                        ev->ixvclo[6 * k + c] = tmp0 * x0[3 * ic + c] + tmp1 * x1[3 * ic + c];
                        ev->ixvclo[6 * k + 3 + c] = tmp0 * v0[3 * ic + c] + tmp1 * v1[3 * ic + c];
                        ev->jxvclo[6 * k + c] = tmp0 * x0[3 * jc + c] + tmp1 * x1[3 * jc + c];
                        ev->jxvclo[6 * k + 3 + c] = tmp0 * v0[3 * jc + c] + tmp1 * v1[3 * jc + c];
                    }
                }
            }
            if (d2min <= d2near) {
                if (ev->nhit >= ev->cap) {
                    ev->overflow = 1;
                } else {
                    int k = ev->nhit++;
                    int ih = i, jh = j;
                    ev->thit[k] = tmin + time;
                    ev->dhit[k] = sqrt(d2min);
                    ev->chit[k] = -1;
                    if (d2min <= d2hit) ev->chit[k] = 1;
                    if (m[jh] > m[ih] && j < nbig) { ih = j; jh = i; }
                    ev->ihit[k] = ih + 1;
                    ev->jhit[k] = jh + 1;
                    if ((tmin - ev->thit1) * h < 0.0) {
                        ev->thit1 = tmin;
                        ev->nowflag = 0;
                        if (d1 <= d2hit) ev->nowflag = 1;
                    }
                }
            }
        }
    }
    free(xmin); free(xmax); free(ymin); free(ymax);
}

/* mce_cent  mercury6_2.for:1059-1185 (algor != 11 branch) */
int orc_mce_cent(double time, double h, double rcen, int i0, int nbod,
                 const double *m, const double *x0, const double *v0,
                 const double *x1, const double *v1, int cap, int *jhit,
                 double *thit, double *dhit)
{
    if (i0 <= 0) i0 = 2;
    int nhit = 0;
    double rcen2 = rcen * rcen;
    double mcen = m[0];
    for (int jj = i0; jj <= nbod; jj++) {
        int j = jj - 1;
        double rr0 = x0[3 * j] * x0[3 * j] + x0[3 * j + 1] * x0[3 * j + 1] + x0[3 * j + 2] * x0[3 * j + 2];
        double rr1 = x1[3 * j] * x1[3 * j] + x1[3 * j + 1] * x1[3 * j + 1] + x1[3 * j + 2] * x1[3 * j + 2];
        double rv0 = v0[3 * j] * x0[3 * j] + v0[3 * j + 1] * x0[3 * j + 1] + v0[3 * j + 2] * x0[3 * j + 2];
        double rv1 = v1[3 * j] * x1[3 * j] + v1[3 * j + 1] * x1[3 * j + 1] + v1[3 * j + 2] * x1[3 * j + 2];
        if ((rv0 * h <= 0.0 && rv1 * h >= 0.0) || fmin(rr0, rr1) <= rcen2) {
            double hx = x0[3 * j + 1] * v0[3 * j + 2] - x0[3 * j + 2] * v0[3 * j + 1];
            double hy = x0[3 * j + 2] * v0[3 * j] - x0[3 * j] * v0[3 * j + 2];
            double hz = x0[3 * j] * v0[3 * j + 1] - x0[3 * j + 1] * v0[3 * j];
            double v2 = v0[3 * j] * v0[3 * j] + v0[3 * j + 1] * v0[3 * j + 1] + v0[3 * j + 2] * v0[3 * j + 2];
            double h2 = hx * hx + hy * hy + hz * hz;
            double p = h2 / (mcen + m[j]);
            double r0 = sqrt(rr0);
            double temp = 1.0 + p * (v2 / (mcen + m[j]) - 2.0 / r0);
            double e = sqrt(fmax(temp, 0.0));
            double q = p / (1.0 + e);
            if (q <= rcen) {
                if (nhit >= cap) return -1;
                int k = nhit++;
                jhit[k] = jj;
                dhit[k] = rcen;
                temp = 0.0;
                if (e > 0.0 && e < 1.0) {
                    double a = q / (1.0 - e);
                    double uhit = fsign(acos(fmax(-1.0, fmin(1.0, (1.0 - rcen / a) / e))), -h);
                    double u0 = fsign(acos(fmax(-1.0, fmin(1.0, (1.0 - r0 / a) / e))), rv0);
                    double mhit = uhit - e * sin(uhit);
                    double m0 = u0 - e * sin(u0);
                    temp = (mhit - m0) / sqrt((mcen + m[j]) / (a * a * a));
                } else if (e > 1.0) {
                    double a = q / (e - 1.0);
                    double uhit = fsign(acosh(fmax(1.0, (1.0 + rcen / a) / e)), -h);
                    double u0 = fsign(acosh(fmax(1.0, (1.0 + r0 / a) / e)), rv0);
                    double mhit = e * sinh(uhit) - uhit;
                    double m0 = e * sinh(u0) - u0;
                    temp = (mhit - m0) / sqrt((mcen + m[j]) / (a * a * a));
                } else if (e == 1.0 && q > 0.0) {
                    double uhit = fsign(sqrt(fmax(0.0, rcen / q - 1.0)), -h);
                    double u0 = fsign(sqrt(fmax(0.0, r0 / q - 1.0)), rv0);
                    temp = sqrt(2.0 * (q * q * q) / (mcen + m[j]))
                         * (uhit + (uhit * uhit * uhit) / 3.0 - u0 - (u0 * u0 * u0) / 3.0);
                }
                if (q == 0.0 && rr1 != rr0)
                    temp = h * fmax(0.0, fmin(1.0, (r0 - rcen) / (r0 - sqrt(rr1))));
                thit[k] = time - h + temp;
            }
        }
    }
    return nhit;
}

/* ------------------------------------------------------------------------ */
/* mdt_mvs  mercury6_2.for:3987-4097 (no user force; cometary forces from the
 * orc_set_ngf model) */
void orc_mdt_mvs(double time, double h0, double rcen, const double *jcen,
                 int nbod, int nbig, const double *m, double *x, double *v,
                 double *a, int *dtflag, const double *rce, const double *rphys,
                 orc_events *ev)
{
    double hby2 = .5 * h0;
    size_t n3 = (size_t)3 * nbod;
    double *xj = calloc(n3, sizeof(double)), *vj = calloc(n3, sizeof(double));
    double *x0 = malloc(n3 * sizeof(double)), *v0 = malloc(n3 * sizeof(double));
    double *gm = calloc((size_t)nbod, sizeof(double));
    if (ev) {
        ev->nclo = 0;
        ev->overflow = 0;
    }
    /* mdt_mvs' cometary accelerations angf (zero without the model) */
    double *angf = (g_angf && (g_ngflag == 1 || g_ngflag == 3)) ? g_angf : NULL;
    double *zero = angf ? NULL : calloc(n3, sizeof(double));
    double *ang = angf ? angf : zero;
    if (*dtflag != 2) {
        *dtflag = 2;
        orc_mco_h2j(nbig, m, x, v, xj, vj);
        orc_mfo_mvs(jcen, nbod, nbig, m, x, xj, a);
        if (angf) orc_mfo_ngf(nbod, x, v, angf, g_ngf);
    }
    double minside = m[0];
    for (int j = 1; j < nbig; j++) {
        double msofar = minside + m[j];
        gm[j] = m[0] * msofar / minside;
        minside = msofar;
    }
    for (int j = 1; j < nbod; j++) {
        v[3 * j + 0] = v[3 * j + 0] + hby2 * (ang[3 * j + 0] + 0.0 + a[3 * j + 0]);
        v[3 * j + 1] = v[3 * j + 1] + hby2 * (ang[3 * j + 1] + 0.0 + a[3 * j + 1]);
        v[3 * j + 2] = v[3 * j + 2] + hby2 * (ang[3 * j + 2] + 0.0 + a[3 * j + 2]);
    }
    memcpy(x0, x, n3 * sizeof(double));
    memcpy(v0, v, n3 * sizeof(double));
    orc_mco_h2j(nbig, m, x, v, xj, vj);
    for (int j = 1; j < nbig; j++) {
        double xv[6] = {xj[3 * j], xj[3 * j + 1], xj[3 * j + 2], vj[3 * j], vj[3 * j + 1], vj[3 * j + 2]};
        orc_drift_one(gm[j], xv, h0);
        for (int c = 0; c < 3; c++) { xj[3 * j + c] = xv[c]; vj[3 * j + c] = xv[3 + c]; }
    }
    for (int j = nbig; j < nbod; j++) {
        double xv[6] = {x[3 * j], x[3 * j + 1], x[3 * j + 2], v[3 * j], v[3 * j + 1], v[3 * j + 2]};
        orc_drift_one(m[0], xv, h0);
        for (int c = 0; c < 3; c++) { x[3 * j + c] = xv[c]; v[3 * j + c] = xv[3 + c]; }
    }
    orc_mco_j2h(nbig, m, xj, vj, x, v);
    if (ev)
        orc_mce_stat(time + h0, h0, rcen, nbod, nbig, m, x0, v0, x, v, rce, rphys, ev);
    orc_mfo_mvs(jcen, nbod, nbig, m, x, xj, a);
    if (angf) orc_mfo_ngf(nbod, x, v, angf, g_ngf);
    for (int j = 1; j < nbod; j++) {
        v[3 * j + 0] = v[3 * j + 0] + hby2 * (ang[3 * j + 0] + 0.0 + a[3 * j + 0]);
        v[3 * j + 1] = v[3 * j + 1] + hby2 * (ang[3 * j + 1] + 0.0 + a[3 * j + 1]);
        v[3 * j + 2] = v[3 * j + 2] + hby2 * (ang[3 * j + 2] + 0.0 + a[3 * j + 2]);
    }
    free(xj); free(vj); free(x0); free(v0); free(gm); free(zero);
}

/* mco_mvs2h  mercury6_2.for:2107-2219 (with mfo_ngf's angf in the kicks when
 * the cometary model is on, 2171-2172 and 2200-2201) */
void orc_mco_mvs2h(double h, const double *jcen, int nbod, int nbig,
                   const double *m, const double *x, const double *v,
                   double *xh, double *vh)
{
    size_t n3 = (size_t)3 * nbod;
    double *xj = calloc(n3, sizeof(double)), *vj = calloc(n3, sizeof(double));
    double *a = calloc(n3, sizeof(double)), *gm = calloc((size_t)nbod, sizeof(double));
    double *ang = calloc(n3, sizeof(double));   /* angf: zero unless the cometary model is on */
    double rt10 = sqrt(10.0);
    double ha[2] = {h * rt10 * 3.0 / 10.0, h * rt10 / 5.0};
    double hb[2] = {-h * rt10 / 72.0, h * rt10 / 24.0};
    memcpy(xh, x, n3 * sizeof(double));
    memcpy(vh, v, n3 * sizeof(double));
    double minside = m[0];
    for (int j = 1; j < nbig; j++) {
        double msofar = minside + m[j];
        gm[j] = m[0] * msofar / minside;
        minside = msofar;
    }
    for (int k = 0; k < 2; k++) {
        double dts[3] = {ha[k], -2.0 * ha[k], ha[k]};
        double kick[2] = {-hb[k], hb[k]};
        for (int phase = 0; phase < 3; phase++) {
            orc_mco_h2j(nbig, m, xh, vh, xj, vj);
            for (int j = 1; j < nbig; j++) {
                double xv[6] = {xj[3 * j], xj[3 * j + 1], xj[3 * j + 2], vj[3 * j], vj[3 * j + 1], vj[3 * j + 2]};
                orc_drift_one(gm[j], xv, dts[phase]);
                for (int c = 0; c < 3; c++) { xj[3 * j + c] = xv[c]; vj[3 * j + c] = xv[3 + c]; }
            }
            for (int j = nbig; j < nbod; j++) {
                double xv[6] = {xh[3 * j], xh[3 * j + 1], xh[3 * j + 2], vh[3 * j], vh[3 * j + 1], vh[3 * j + 2]};
                orc_drift_one(m[0], xv, dts[phase]);
                for (int c = 0; c < 3; c++) { xh[3 * j + c] = xv[c]; vh[3 * j + c] = xv[3 + c]; }
            }
            orc_mco_j2h(nbig, m, xj, vj, xh, vh);
            if (phase == 2) break;
            orc_mfo_mvs(jcen, nbod, nbig, m, xh, xj, a);
            if (g_ngflag == 1 || g_ngflag == 3) orc_mfo_ngf(nbod, xh, vh, ang, g_ngf);
            /* vh - hb*(...) for the first kick, vh + hb*(...) for the second */
            for (int j = 1; j < nbod; j++)
                for (int c = 0; c < 3; c++) {
                    if (phase == 0)
                        vh[3 * j + c] = vh[3 * j + c] - hb[k] * (ang[3 * j + c] + 0.0 + a[3 * j + c]);
                    else
                        vh[3 * j + c] = vh[3 * j + c] + hb[k] * (ang[3 * j + c] + 0.0 + a[3 * j + c]);
                }
            (void)kick;
        }
    }
    free(xj); free(vj); free(a); free(gm); free(ang);
}

/* ------------------------------------------------------------------------ */
int orc_mvs_run(double time0, double h0, double rcen, const double *jcen,
                int nbod, int nbig, const double *m, double *x, double *v,
                const double *rce, const double *rphys, int nsteps,
                int screen, int cap, int *evi, double *evd)
{
    size_t n3 = (size_t)3 * nbod;
    double *a = calloc(n3, sizeof(double));
    double *xs = malloc(n3 * sizeof(double)), *vs = malloc(n3 * sizeof(double));
    int lcap = 2 * nbod > 4096 ? 2 * nbod : 4096;   /* events of one step */
    orc_events ev;
    memset(&ev, 0, sizeof ev);
    ev.cap = lcap;
    ev.iclo = malloc(sizeof(int) * lcap); ev.jclo = malloc(sizeof(int) * lcap);
    ev.dclo = malloc(sizeof(double) * lcap); ev.tclo = malloc(sizeof(double) * lcap);
    ev.ixvclo = malloc(sizeof(double) * 6 * lcap); ev.jxvclo = malloc(sizeof(double) * 6 * lcap);
    ev.ihit = malloc(sizeof(int) * lcap); ev.jhit = malloc(sizeof(int) * lcap);
    ev.chit = malloc(sizeof(int) * lcap); ev.dhit = malloc(sizeof(double) * lcap);
    ev.thit = malloc(sizeof(double) * lcap);
    int *cj = malloc(sizeof(int) * nbod);
    double *ct = malloc(sizeof(double) * nbod), *cd = malloc(sizeof(double) * nbod);
    int dtflag = 0, nev = 0, bad = 0;
    double time = time0;
    for (int s = 0; s < nsteps; s++) {
        memcpy(xs, x, n3 * sizeof(double));
        memcpy(vs, v, n3 * sizeof(double));
        orc_mdt_mvs(time, h0, rcen, jcen, nbod, nbig, m, x, v, a, &dtflag, rce, rphys,
                    screen ? &ev : NULL);
        time = time + h0;
        if (!screen) continue;
        if (ev.overflow) bad = 1;
        for (int k = 0; k < ev.nclo; k++) {
            if (nev >= cap) { bad = 1; break; }
            evi[4 * nev] = s; evi[4 * nev + 1] = 1;
            evi[4 * nev + 2] = ev.iclo[k]; evi[4 * nev + 3] = ev.jclo[k];
            evd[2 * nev] = ev.dclo[k]; evd[2 * nev + 1] = ev.tclo[k];
            nev++;
        }
        for (int k = 0; k < ev.nhit; k++) {
            if (nev >= cap) { bad = 1; break; }
            evi[4 * nev] = s; evi[4 * nev + 1] = ev.chit[k] > 0 ? 6 : 2;
            evi[4 * nev + 2] = ev.ihit[k]; evi[4 * nev + 3] = ev.jhit[k];
            evd[2 * nev] = ev.dhit[k]; evd[2 * nev + 1] = ev.thit[k];
            nev++;
        }
        int nc = orc_mce_cent(time, h0, rcen, 2, nbod, m, xs, vs, x, v, nbod, cj, ct, cd);
        for (int k = 0; k < nc; k++) {
            if (nev >= cap) { bad = 1; break; }
            evi[4 * nev] = s; evi[4 * nev + 1] = 8;
            evi[4 * nev + 2] = 1; evi[4 * nev + 3] = cj[k];
            evd[2 * nev] = cd[k]; evd[2 * nev + 1] = ct[k];
            nev++;
        }
    }
    free(a); free(xs); free(vs); free(cj); free(ct); free(cd);
    free(ev.iclo); free(ev.jclo); free(ev.dclo); free(ev.tclo); free(ev.ixvclo); free(ev.jxvclo);
    free(ev.ihit); free(ev.jhit); free(ev.chit); free(ev.dhit); free(ev.thit);
    return bad ? -1 : nev;
}

/* ------------------------------------------------------------------------ */
/* mfo_grav  mercury6_2.for:4593-4707 */
static void mfo_grav(int nbod, int nbig, const double *m, const double *x, const double *v,
                     double *a, int pn)
{
    double sx = 0.0, sy = 0.0, sz = 0.0;
    double *r3 = malloc(sizeof(double) * (size_t)(nbod > 0 ? nbod : 1));
    double *ri = malloc(sizeof(double) * (size_t)(nbod > 0 ? nbod : 1));
    for (int i = 1; i < nbod; i++) {
        a[3 * i] = a[3 * i + 1] = a[3 * i + 2] = 0.0;
        double s2 = x[3 * i] * x[3 * i] + x[3 * i + 1] * x[3 * i + 1] + x[3 * i + 2] * x[3 * i + 2];
        double s_1 = 1.0 / sqrt(s2);
        ri[i] = s_1;
        r3[i] = s_1 * s_1 * s_1;
    }
    for (int i = 1; i < nbod; i++) {
        double tmp1 = m[i] * r3[i];
        sx = sx - tmp1 * x[3 * i];
        sy = sy - tmp1 * x[3 * i + 1];
        sz = sz - tmp1 * x[3 * i + 2];
    }
    for (int i = 1; i < nbig; i++) {
        for (int j = i + 1; j < nbod; j++) {
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double s2 = dx * dx + dy * dy + dz * dz;
            double s_1 = 1.0 / sqrt(s2);
            double s_3 = s_1 * s_1 * s_1;
            double tmp1 = s_3 * m[i];
            double tmp2 = s_3 * m[j];
            a[3 * j] = a[3 * j] - tmp1 * dx;
            a[3 * j + 1] = a[3 * j + 1] - tmp1 * dy;
            a[3 * j + 2] = a[3 * j + 2] - tmp1 * dz;
            a[3 * i] = a[3 * i] + tmp2 * dx;
            a[3 * i + 1] = a[3 * i + 1] + tmp2 * dy;
            a[3 * i + 2] = a[3 * i + 2] + tmp2 * dz;
        }
    }
    if (pn == 0) {
        for (int i = 1; i < nbod; i++) {
            double tmp1 = m[0] * r3[i];
            a[3 * i] = a[3 * i] + sx - tmp1 * x[3 * i];
            a[3 * i + 1] = a[3 * i + 1] + sy - tmp1 * x[3 * i + 1];
            a[3 * i + 2] = a[3 * i + 2] + sz - tmp1 * x[3 * i + 2];
        }
    } else {
        /* (c in AU/day)**2; AU*.01d0 is the AU in metres */
        const double cl = 299792458.0 * 86400.0 / (1.495978707e13 * .01);
        const double light2 = cl * cl;
        for (int i = 1; i < nbod; i++) {
            const double *xi = x + 3 * i, *vi = v + 3 * i;
            double tmp1 = m[0] * r3[i];
            double rv = xi[0] * vi[0] + xi[1] * vi[1] + xi[2] * vi[2];
            double v2 = vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2];
            double f = tmp1 / light2;
            double radial = 4.0 * m[0] * ri[i] - v2;
            a[3 * i] = a[3 * i] + sx - tmp1 * xi[0] + f * (radial * xi[0] + 4.0 * rv * vi[0]);
            a[3 * i + 1] = a[3 * i + 1] + sy - tmp1 * xi[1] + f * (radial * xi[1] + 4.0 * rv * vi[1]);
            a[3 * i + 2] = a[3 * i + 2] + sz - tmp1 * xi[2] + f * (radial * xi[2] + 4.0 * rv * vi[2]);
        }
    }
    free(r3);
    free(ri);
}

/* The non-gravitational force model of the calls that follow (the driver's
 * ngflag and ngf(5,nbod)): ngflag 1 cometary A1..A3 and Yarkovsky, 2 radiation
 * pressure and Poynting-Robertson drag, 3 both; 0 none.  It stays set until
 * changed; orc_set_ngf(0, NULL, 0) clears it.  The MVS step's cometary
 * accelerations (mdt_mvs' angf) are carried in the model, like mdt_mvs'
 * saved array. */
void orc_set_ngf(int ngflag, const double *ngf, int nbod)
{
    g_ngflag = ngf ? ngflag : 0;
    g_ngf = ngf;
    free(g_angf);
    g_angf = (ngf && nbod > 0) ? calloc((size_t)3 * nbod, sizeof(double)) : NULL;
}

/* mfo_ngf  mercury6_2.for:5106-5152: cometary A1/A2/A3 with Marsden's g(r)
 * and the distance cutoff, and the Yarkovsky transverse term.  Returns 1 for
 * an undefined transverse or normal direction (the Fortran's fail). */
int orc_mfo_ngf(int nbod, const double *x, const double *v, double *a, const double *ngf)
{
    for (int k = 0; k < 3 * nbod; k++) a[k] = 0.0;
    for (int j = 1; j < nbod; j++) {
        const double *p = ngf + 5 * j;
        const double *xj = x + 3 * j, *vj = v + 3 * j;
        double *aj = a + 3 * j;
        if (p[0] == 0.0 && p[1] == 0.0 && p[2] == 0.0 && p[4] == 0.0) continue;
        double r2 = xj[0] * xj[0] + xj[1] * xj[1] + xj[2] * xj[2];
        double r = sqrt(r2);
        double rv = (xj[0] * vj[0] + xj[1] * vj[1] + xj[2] * vj[2]) / r2;
        double tx = vj[0] - rv * xj[0];
        double ty = vj[1] - rv * xj[1];
        double tz = vj[2] - rv * xj[2];
        double tn = sqrt(tx * tx + ty * ty + tz * tz);
        if ((p[0] != 0.0 || p[1] != 0.0 || p[2] != 0.0) &&
            (r2 < 88.0 || fabs(p[0]) > 1e-7 || fabs(p[1]) > 1e-7 || fabs(p[2]) > 1e-7)) {
            double q = r * .3561253561253561;
            double g = .111262 * pow(q, -2.15) * pow(1.0 + pow(q, 5.093), -4.6142);
            for (int c = 0; c < 3; c++) aj[c] = p[0] * g / r * xj[c];
            if (p[1] != 0.0) {
                if (tn == 0.0) return 1;
                double f = p[1] * g / tn;
                aj[0] = aj[0] + f * tx; aj[1] = aj[1] + f * ty; aj[2] = aj[2] + f * tz;
            }
            if (p[2] != 0.0) {
                double nx = xj[1] * vj[2] - xj[2] * vj[1];
                double ny = xj[2] * vj[0] - xj[0] * vj[2];
                double nz = xj[0] * vj[1] - xj[1] * vj[0];
                double nn = sqrt(nx * nx + ny * ny + nz * nz);
                if (nn == 0.0) return 1;
                double f = p[2] * g / nn;
                aj[0] = aj[0] + f * nx; aj[1] = aj[1] + f * ny; aj[2] = aj[2] + f * nz;
            }
        }
        if (p[4] != 0.0) {
            if (tn == 0.0) return 1;
            double f = p[4] / (r2 * tn);
            aj[0] = aj[0] + f * tx; aj[1] = aj[1] + f * ty; aj[2] = aj[2] + f * tz;
        }
    }
    return 0;
}

/* mfo_pr  mercury6_2.for:5307-5357: radiation pressure and Poynting-Robertson
 * drag (solar-wind factor 0.3) on massless bodies with beta = ngf(4). */
void orc_mfo_pr(int nbod, const double *m, const double *x, const double *v, double *a, const double *ngf)
{
    const double sw = .3;
    const double c = 299792458.0 * 86400.0 / (1.495978707e13 * .01);
    for (int j = 1; j < nbod; j++) {
        const double *xj = x + 3 * j, *vj = v + 3 * j;
        double *aj = a + 3 * j;
        double beta = ngf[5 * j + 3];
        if (m[j] != 0.0 || beta == 0.0) {
            aj[0] = 0.0; aj[1] = 0.0; aj[2] = 0.0;
            continue;
        }
        double r = sqrt(xj[0] * xj[0] + xj[1] * xj[1] + xj[2] * xj[2]);
        double vr = (xj[0] * vj[0] + xj[1] * vj[1] + xj[2] * vj[2]) / r;
        double f = K2 * beta / (r * r);
        for (int k = 0; k < 3; k++)
            aj[k] = f * ((1.0 - (1.0 + sw) * vr / c) * xj[k] / r - (1.0 + sw) * vj[k] / c);
    }
}

/* mfo_all  mercury6_2.for:4506-4592 (gravity, 1PN, oblateness and the
 * non-gravitational model of orc_set_ngf).  Returns 1 when mfo_ngf fails. */
static int mfo_all_impl(const double *jcen, int nbod, int nbig, const double *m,
                        const double *x, const double *v, int pn, double *a)
{
    mfo_grav(nbod, nbig, m, x, v, a, pn);
    double *acor = NULL;
    if (jcen[0] != 0 || jcen[1] != 0 || jcen[2] != 0) {
        acor = calloc((size_t)3 * nbod, sizeof(double));
        double acen[3];
        orc_mfo_obl(jcen, nbod, m, x, acor, acen);
        for (int j = 1; j < nbod; j++)
            for (int c = 0; c < 3; c++)
                a[3 * j + c] = a[3 * j + c] + (acor[3 * j + c] - acen[c]);
    }
    int bad = 0;
    if (g_ngflag == 1 || g_ngflag == 3) {
        if (!acor) acor = calloc((size_t)3 * nbod, sizeof(double));
        bad = orc_mfo_ngf(nbod, x, v, acor, g_ngf);
        for (int j = 1; j < nbod; j++)
            for (int c = 0; c < 3; c++) a[3 * j + c] = a[3 * j + c] + acor[3 * j + c];
    }
    if (g_ngflag == 2 || g_ngflag == 3) {
        if (!acor) acor = calloc((size_t)3 * nbod, sizeof(double));
        orc_mfo_pr(nbod, m, x, v, acor, g_ngf);
        for (int j = 1; j < nbod; j++)
            for (int c = 0; c < 3; c++) a[3 * j + c] = a[3 * j + c] + acor[3 * j + c];
    }
    free(acor);
    return bad;
}

void orc_mfo_all(const double *jcen, int nbod, int nbig, const double *m,
                 const double *x, const double *v, int pn, double *a)
{
    (void)mfo_all_impl(jcen, nbod, nbig, m, x, v, pn, a);
}

/* mdt_ra15 constants (mercury6_2.for:4162-4199) */
static const double RA_H[8] = {0.0, .0562625605369221, .1802406917368924, .3526247171131696,
                               .5471536263305554, .7342101772154105, .8853209468390958,
                               .9775206135612875};
static const double RA_XC[8] = {.5, .1666666666666667, .08333333333333333, .05, .03333333333333333,
                                .02380952380952381, .01785714285714286, .01388888888888889};
static const double RA_VC[7] = {.5, .3333333333333333, .25, .2, .1666666666666667,
                                .1428571428571429, .125};

static void ra_constants(double *r, double *c, double *d)
{
    const double *h = RA_H;
    int n = 0;
    for (int j = 1; j < 8; j++)
        for (int k = 0; k < j; k++)
            r[n++] = 1.0 / (h[j] - h[k]);
    /* c, d: Fortran 1-based recurrences, stored 0-based */
    c[0] = -h[1];
    d[0] = h[1];
    n = 1;   /* Fortran n */
    for (int j = 3; j <= 7; j++) {
        n = n + 1;
        c[n - 1] = -h[j - 1] * c[n - j + 2 - 1];
        d[n - 1] = h[1] * d[n - j + 2 - 1];
        for (int k = 3; k <= j - 1; k++) {
            n = n + 1;
            c[n - 1] = c[n - j + 1 - 1] - h[j - 1] * c[n - j + 2 - 1];
            d[n - 1] = d[n - j + 1 - 1] + h[k - 1] * d[n - j + 2 - 1];
        }
        n = n + 1;
        c[n - 1] = c[n - j + 1 - 1] - h[j - 1];
        d[n - 1] = d[n - j + 1 - 1] + h[j - 1];
    }
}

/* mdt_ra15  mercury6_2.for:4133-4500 */
long orc_mdt_ra15(double time, double *t, double *tdid, double tol, const double *jcen,
                  int nbod, int nbig, const double *m, double *x1, double *v1, int pn,
                  orc_ra15 *st, int *redo)
{
    (void)time;
    const double *h = RA_H;
    double r[28], c[21], d[21], s[9];
    ra_constants(r, c, d);
    if (st->dtflag == 0) st->dtflag = 1;
    const int nv = 3 * nbod;
    double *x = calloc((size_t)nv, sizeof(double)), *v = calloc((size_t)nv, sizeof(double));
    double *a = calloc((size_t)nv, sizeof(double)), *a1 = calloc((size_t)nv, sizeof(double));
    double *g = calloc((size_t)7 * nv, sizeof(double));
    double *b = st->b, *e = st->e;
    long nforce = 0;
    *redo = 0;
    for (;;) {
        int niter;
        if (st->dtflag == 1) {
            niter = 6;
            for (int k = 3; k < nv; k++)
                for (int q = 0; q < 7; q++) { b[7 * k + q] = 0.0; e[7 * k + q] = 0.0; }
        } else {
            niter = 2;
        }
        orc_mfo_all(jcen, nbod, nbig, m, x1, v1, pn, a1);
        nforce++;
        for (int k = 3; k < nv; k++) {
            double *bk = b + 7 * k, *gk = g + 7 * k;
            gk[0] = bk[6] * d[15] + bk[5] * d[10] + bk[4] * d[6] + bk[3] * d[3] + bk[2] * d[1] + bk[1] * d[0] + bk[0];
            gk[1] = bk[6] * d[16] + bk[5] * d[11] + bk[4] * d[7] + bk[3] * d[4] + bk[2] * d[2] + bk[1];
            gk[2] = bk[6] * d[17] + bk[5] * d[12] + bk[4] * d[8] + bk[3] * d[5] + bk[2];
            gk[3] = bk[6] * d[18] + bk[5] * d[13] + bk[4] * d[9] + bk[3];
            gk[4] = bk[6] * d[19] + bk[5] * d[14] + bk[4];
            gk[5] = bk[6] * d[20] + bk[5];
            gk[6] = bk[6];
        }
        const int vpred = pn != 0 || g_ngflag != 0;   /* mercury6_2.for:4277 */
        for (int n = 1; n <= niter; n++) {
            for (int j = 2; j <= 8; j++) {
                double hj = h[j - 1];
                s[0] = *t * hj;
                s[1] = s[0] * s[0] * .5;
                s[2] = s[1] * hj * .3333333333333333;
                s[3] = s[2] * hj * .5;
                s[4] = s[3] * hj * .6;
                s[5] = s[4] * hj * .6666666666666667;
                s[6] = s[5] * hj * .7142857142857143;
                s[7] = s[6] * hj * .75;
                s[8] = s[7] * hj * .7777777777777778;
                for (int k = 3; k < nv; k++) {
                    const double *bk = b + 7 * k;
                    x[k] = s[8] * bk[6] + s[7] * bk[5] + s[6] * bk[4] + s[5] * bk[3] + s[4] * bk[2]
                         + s[3] * bk[1] + s[2] * bk[0] + s[1] * a1[k] + s[0] * v1[k] + x1[k];
                }
                if (vpred) {
                    s[0] = *t * hj;
                    s[1] = s[0] * hj * .5;
                    s[2] = s[1] * hj * .6666666666666667;
                    s[3] = s[2] * hj * .75;
                    s[4] = s[3] * hj * .8;
                    s[5] = s[4] * hj * .8333333333333333;
                    s[6] = s[5] * hj * .8571428571428571;
                    s[7] = s[6] * hj * .875;
                    for (int k = 3; k < nv; k++) {
                        const double *bk = b + 7 * k;
                        v[k] = s[7] * bk[6] + s[6] * bk[5] + s[5] * bk[4] + s[4] * bk[3] + s[3] * bk[2]
                             + s[2] * bk[1] + s[1] * bk[0] + s[0] * a1[k] + v1[k];
                    }
                }
                orc_mfo_all(jcen, nbod, nbig, m, x, v, pn, a);
                nforce++;
                for (int k = 3; k < nv; k++) {
                    double *bk = b + 7 * k, *gk = g + 7 * k;
                    double temp, gkk = a[k] - a1[k];
                    switch (j) {
                    case 2:
                        temp = gk[0];
                        gk[0] = (a[k] - a1[k]) * r[0];
                        bk[0] = bk[0] + gk[0] - temp;
                        break;
                    case 3:
                        temp = gk[1];
                        gk[1] = (gkk * r[1] - gk[0]) * r[2];
                        temp = gk[1] - temp;
                        bk[0] = bk[0] + temp * c[0];
                        bk[1] = bk[1] + temp;
                        break;
                    case 4:
                        temp = gk[2];
                        gk[2] = ((gkk * r[3] - gk[0]) * r[4] - gk[1]) * r[5];
                        temp = gk[2] - temp;
                        bk[0] = bk[0] + temp * c[1];
                        bk[1] = bk[1] + temp * c[2];
                        bk[2] = bk[2] + temp;
                        break;
                    case 5:
                        temp = gk[3];
                        gk[3] = (((gkk * r[6] - gk[0]) * r[7] - gk[1]) * r[8] - gk[2]) * r[9];
                        temp = gk[3] - temp;
                        bk[0] = bk[0] + temp * c[3];
                        bk[1] = bk[1] + temp * c[4];
                        bk[2] = bk[2] + temp * c[5];
                        bk[3] = bk[3] + temp;
                        break;
                    case 6:
                        temp = gk[4];
                        gk[4] = ((((gkk * r[10] - gk[0]) * r[11] - gk[1]) * r[12] - gk[2]) * r[13] - gk[3]) * r[14];
                        temp = gk[4] - temp;
                        bk[0] = bk[0] + temp * c[6];
                        bk[1] = bk[1] + temp * c[7];
                        bk[2] = bk[2] + temp * c[8];
                        bk[3] = bk[3] + temp * c[9];
                        bk[4] = bk[4] + temp;
                        break;
                    case 7:
                        temp = gk[5];
                        gk[5] = (((((gkk * r[15] - gk[0]) * r[16] - gk[1]) * r[17] - gk[2]) * r[18] - gk[3]) * r[19]
                                 - gk[4]) * r[20];
                        temp = gk[5] - temp;
                        bk[0] = bk[0] + temp * c[10];
                        bk[1] = bk[1] + temp * c[11];
                        bk[2] = bk[2] + temp * c[12];
                        bk[3] = bk[3] + temp * c[13];
                        bk[4] = bk[4] + temp * c[14];
                        bk[5] = bk[5] + temp;
                        break;
                    default:
                        temp = gk[6];
                        gk[6] = ((((((gkk * r[21] - gk[0]) * r[22] - gk[1]) * r[23] - gk[2]) * r[24] - gk[3]) * r[25]
                                  - gk[4]) * r[26] - gk[5]) * r[27];
                        temp = gk[6] - temp;
                        bk[0] = bk[0] + temp * c[15];
                        bk[1] = bk[1] + temp * c[16];
                        bk[2] = bk[2] + temp * c[17];
                        bk[3] = bk[3] + temp * c[18];
                        bk[4] = bk[4] + temp * c[19];
                        bk[5] = bk[5] + temp * c[20];
                        bk[6] = bk[6] + temp;
                        break;
                    }
                }
            }
        }
        double temp = 0.0;
        for (int k = 3; k < nv; k++) temp = fmax(temp, fabs(b[7 * k + 6]));
        temp = temp / (72.0 * pow(fabs(*t), 7.0));
        *tdid = *t;
        if (temp == 0.0)
            *t = *tdid * 1.4;
        else
            *t = fsign(pow(tol / temp, 1.0 / 9.0), *tdid);
        if (st->dtflag == 1 && fabs(*t / *tdid) < 1.0) {
            *t = *t * .8;
            (*redo)++;
            continue;
        }
        if (fabs(*t / *tdid) > 1.4) *t = *tdid * 1.4;
        break;
    }
    double temp = *tdid * *tdid;
    for (int k = 3; k < nv; k++) {
        const double *bk = b + 7 * k;
        x1[k] = (RA_XC[7] * bk[6] + RA_XC[6] * bk[5] + RA_XC[5] * bk[4] + RA_XC[4] * bk[3] + RA_XC[3] * bk[2]
                 + RA_XC[2] * bk[1] + RA_XC[1] * bk[0] + RA_XC[0] * a1[k]) * temp + v1[k] * *tdid + x1[k];
        v1[k] = (RA_VC[6] * bk[6] + RA_VC[5] * bk[5] + RA_VC[4] * bk[4] + RA_VC[3] * bk[3] + RA_VC[2] * bk[2]
                 + RA_VC[1] * bk[1] + RA_VC[0] * bk[0] + a1[k]) * *tdid + v1[k];
    }
    double q = *t / *tdid;
    double q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q3 * q4;
    for (int k = 3; k < nv; k++) {
        double *bk = b + 7 * k, *ek = e + 7 * k;
        double sd[7];
        for (int i = 0; i < 7; i++) sd[i] = bk[i] - ek[i];
        if (st->dtflag == 1)
            for (int i = 0; i < 7; i++) sd[i] = 0.0;
        ek[0] = q * (bk[6] * 7.0 + bk[5] * 6.0 + bk[4] * 5.0 + bk[3] * 4.0 + bk[2] * 3.0 + bk[1] * 2.0 + bk[0]);
        ek[1] = q2 * (bk[6] * 21.0 + bk[5] * 15.0 + bk[4] * 10.0 + bk[3] * 6.0 + bk[2] * 3.0 + bk[1]);
        ek[2] = q3 * (bk[6] * 35.0 + bk[5] * 20.0 + bk[4] * 10.0 + bk[3] * 4.0 + bk[2]);
        ek[3] = q4 * (bk[6] * 35.0 + bk[5] * 15.0 + bk[4] * 5.0 + bk[3]);
        ek[4] = q5 * (bk[6] * 21.0 + bk[5] * 6.0 + bk[4]);
        ek[5] = q6 * (bk[6] * 7.0 + bk[5]);
        ek[6] = q7 * bk[6];
        for (int i = 0; i < 7; i++) bk[i] = ek[i] + sd[i];
    }
    st->dtflag = 2;
    free(x); free(v); free(a); free(a1); free(g);
    return nforce;
}

long orc_ra15_run(double time0, double h0, double tol, const double *jcen, int nbod,
                  int nbig, const double *m, double *x, double *v, int pn, int nsteps,
                  double *hdid, double *hnext)
{
    orc_ra15 st;
    st.nbod = nbod;
    st.dtflag = 0;
    st.b = calloc((size_t)21 * nbod, sizeof(double));
    st.e = calloc((size_t)21 * nbod, sizeof(double));
    double t = h0, time = time0;
    long nf = 0;
    for (int s = 0; s < nsteps; s++) {
        double tdid;
        int redo;
        nf += orc_mdt_ra15(time, &t, &tdid, tol, jcen, nbod, nbig, m, x, v, pn, &st, &redo);
        hdid[s] = tdid;
        time = time + tdid;
    }
    *hnext = t;
    free(st.b);
    free(st.e);
    return nf;
}

/* ------------------------------------------------------------------------ */
/* HYBRID (algorithm 10) */

/* mfo_drct  mercury6_2.for:4708-4780 (i0 = 2, changeover weighting K) */
static void mfo_drct(int nbod, int nbig, const double *m, const double *x, const double *rcrit, double *a)
{
    for (int i = 1; i < nbig; i++) {
        for (int j = i + 1; j < nbod; j++) {
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double s2 = dx * dx + dy * dy + dz * dz;
            double rc = fmax(rcrit[i], rcrit[j]);
            double rc2 = rc * rc;
            double tmp2;
            if (s2 >= rc2) {
                double s_1 = 1.0 / sqrt(s2);
                tmp2 = s_1 * s_1 * s_1;
            } else if (s2 <= 0.01 * rc2) {
                tmp2 = 0.0;
            } else {
                double s_1 = 1.0 / sqrt(s2);
                double s = 1.0 / s_1;
                double s_3 = s_1 * s_1 * s_1;
                double q = (s - 0.1 * rc) / (0.9 * rc);
                double q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3;
                tmp2 = (10.0 * q3 - 15.0 * q4 + 6.0 * q5) * s_3;
            }
            double faci = tmp2 * m[i];
            double facj = tmp2 * m[j];
            a[3 * j] = a[3 * j] - faci * dx;
            a[3 * j + 1] = a[3 * j + 1] - faci * dy;
            a[3 * j + 2] = a[3 * j + 2] - faci * dz;
            a[3 * i] = a[3 * i] + facj * dx;
            a[3 * i + 1] = a[3 * i + 1] + facj * dy;
            a[3 * i + 2] = a[3 * i + 2] + facj * dz;
        }
    }
}

/* mfo_hy  mercury6_2.for:4781-4835 */
void orc_mfo_hy(const double *jcen, int nbod, int nbig, const double *m, const double *x,
                const double *rcrit, double *a)
{
    for (int k = 0; k < 3 * nbod; k++) a[k] = 0.0;
    mfo_drct(nbod, nbig, m, x, rcrit, a);
    if (jcen[0] != 0 || jcen[1] != 0 || jcen[2] != 0) {
        double *aobl = calloc((size_t)3 * nbod, sizeof(double));
        double acen[3];
        orc_mfo_obl(jcen, nbod, m, x, aobl, acen);
        for (int k = 1; k < nbod; k++)
            for (int c = 0; c < 3; c++) a[3 * k + c] = a[3 * k + c] + aobl[3 * k + c] - acen[c];
        free(aobl);
    }
}

/* mce_snif  mercury6_2.for:1648-1764 (i0 = 2).  ice/jce 1-based, in loop
 * order; returns nce, or -1 when more than cap pairs would be listed. */
int orc_mce_snif(double h, int nbod, int nbig, const double *x0, const double *v0, const double *x1,
                 const double *v1, const double *rcrit, int *ce, int cap, int *ice, int *jce)
{
    int nce = 0, over = 0;
    for (int j = 1; j < nbod; j++) ce[j] = 0;
    double *xmin = malloc(sizeof(double) * (size_t)nbod), *xmax = malloc(sizeof(double) * (size_t)nbod);
    double *ymin = malloc(sizeof(double) * (size_t)nbod), *ymax = malloc(sizeof(double) * (size_t)nbod);
    orc_mce_box(nbod, h, x0, v0, x1, v1, xmin, xmax, ymin, ymax);
    for (int j = 1; j < nbig; j++) {
        xmin[j] = xmin[j] - rcrit[j];
        xmax[j] = xmax[j] + rcrit[j];
        ymin[j] = ymin[j] - rcrit[j];
        ymax[j] = ymax[j] + rcrit[j];
    }
    for (int i = 1; i < nbig; i++) {
        for (int j = i + 1; j < nbod; j++) {
            if (!(xmax[i] >= xmin[j] && xmax[j] >= xmin[i] && ymax[i] >= ymin[j] && ymax[j] >= ymin[i]))
                continue;
            double rc = fmax(rcrit[i], rcrit[j]);
            double rc2 = rc * rc;
            double dx0 = x0[3 * i] - x0[3 * j], dy0 = x0[3 * i + 1] - x0[3 * j + 1], dz0 = x0[3 * i + 2] - x0[3 * j + 2];
            double dx1 = x1[3 * i] - x1[3 * j], dy1 = x1[3 * i + 1] - x1[3 * j + 1], dz1 = x1[3 * i + 2] - x1[3 * j + 2];
            double d0 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0;
            double d1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
            double du0 = v0[3 * i] - v0[3 * j], dv0 = v0[3 * i + 1] - v0[3 * j + 1], dw0 = v0[3 * i + 2] - v0[3 * j + 2];
            double du1 = v1[3 * i] - v1[3 * j], dv1 = v1[3 * i + 1] - v1[3 * j + 1], dw1 = v1[3 * i + 2] - v1[3 * j + 2];
            double d0t = (dx0 * du0 + dy0 * dv0 + dz0 * dw0) * 2.0;
            double d1t = (dx1 * du1 + dy1 * dv1 + dz1 * dw1) * 2.0;
            double d2min = HUGEV, tmin;
            if (d0t * h <= 0 && d1t * h >= 0) orc_mce_min(d0, d1, d0t, d1t, h, &d2min, &tmin);
            double temp = fmin(fmin(d0, d1), d2min);
            if (temp <= rc2) {
                ce[i] = 2;
                ce[j] = 2;
                if (nce >= cap) { over = 1; continue; }
                ice[nce] = i + 1;
                jce[nce] = j + 1;
                nce++;
            }
        }
    }
    free(xmin); free(xmax); free(ymin); free(ymax);
    return over ? -1 : nce;
}

/* mfo_hkce  mercury6_2.for (Keplerian part with close-encounter terms):
 * pair terms weighted 1 - K over the listed pairs, then the solar terms. */
static void mfo_hkce(int nbod, const double *m, const double *x, const double *rcrit, double *a,
                     int nce, const int *ice, const int *jce)
{
    for (int k = 0; k < 3 * nbod; k++) a[k] = 0.0;
    for (int k = 0; k < nce; k++) {
        int i = ice[k] - 1, j = jce[k] - 1;
        double dx = x[3 * j] - x[3 * i];
        double dy = x[3 * j + 1] - x[3 * i + 1];
        double dz = x[3 * j + 2] - x[3 * i + 2];
        double s2 = dx * dx + dy * dy + dz * dz;
        double rc = fmax(rcrit[i], rcrit[j]);
        double rc2 = rc * rc;
        if (s2 < rc2) {
            double s_1 = 1.0 / sqrt(s2);
            double s_3 = s_1 * s_1 * s_1;
            double tmp2;
            if (s2 <= 0.01 * rc2) {
                tmp2 = s_3;
            } else {
                double s = 1.0 / s_1;
                double q = (s - 0.1 * rc) / (0.9 * rc);
                double q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3;
                tmp2 = (1.0 - 10.0 * q3 + 15.0 * q4 - 6.0 * q5) * s_3;
            }
            double faci = tmp2 * m[i];
            double facj = tmp2 * m[j];
            a[3 * j] = a[3 * j] - faci * dx;
            a[3 * j + 1] = a[3 * j + 1] - faci * dy;
            a[3 * j + 2] = a[3 * j + 2] - faci * dz;
            a[3 * i] = a[3 * i] + facj * dx;
            a[3 * i + 1] = a[3 * i + 1] + facj * dy;
            a[3 * i + 2] = a[3 * i + 2] + facj * dz;
        }
    }
    for (int i = 1; i < nbod; i++) {
        double s2 = x[3 * i] * x[3 * i] + x[3 * i + 1] * x[3 * i + 1] + x[3 * i + 2] * x[3 * i + 2];
        double s_1 = 1.0 / sqrt(s2);
        double tmp2 = m[0] * s_1 * s_1 * s_1;
        a[3 * i] = a[3 * i] - tmp2 * x[3 * i];
        a[3 * i + 1] = a[3 * i + 1] - tmp2 * x[3 * i + 1];
        a[3 * i + 2] = a[3 * i + 2] - tmp2 * x[3 * i + 2];
    }
}

void orc_mfo_hkce(int nbod, const double *m, const double *x, const double *rcrit, double *a, int nce,
                  const int *ice, const int *jce)
{
    mfo_hkce(nbod, m, x, rcrit, a, nce, ice, jce);
}

/* Force of a BS step: a[3*nbod] at positions x (and velocities v). */
typedef void (*bs_force_fn)(void *ctx, const double *x, const double *v, double *a);

struct hkce_ctx { int nbod; const double *m, *rcrit; int nce; const int *ice, *jce; };
static void hkce_force(void *p, const double *x, const double *v, double *a)
{
    (void)v;
    struct hkce_ctx *c = (struct hkce_ctx *)p;
    mfo_hkce(c->nbod, c->m, x, c->rcrit, a, c->nce, c->ice, c->jce);
}

struct all_ctx { const double *jcen; int nbod, nbig; const double *m; int pn; };
static void all_force(void *p, const double *x, const double *v, double *a)
{
    struct all_ctx *c = (struct all_ctx *)p;
    orc_mfo_all(c->jcen, c->nbod, c->nbig, c->m, x, v, c->pn, a);
}

/* mdt_bs2  mercury6_2.for:3428-3617: 12 modified-midpoint levels of a
 * conservative system (every force call gets the start velocities v0).
 * Returns the number of halvings (rejected attempts); -1 if time cannot
 * advance, -2 on a non-finite extrapolation. */
static int bs2_core(double time, double *h0, double *hdid, double tol, int nbod, double *x0, double *v0,
                    bs_force_fn force, void *fc, long *nforce)
{
    const double SHRINK = .55, GROW = 1.3;
    const double tol2 = tol * tol;
    const size_t nv = (size_t)3 * nbod;
    double *xend = calloc(nv, sizeof(double)), *b = calloc(nv, sizeof(double)), *c = calloc(nv, sizeof(double));
    double *a = calloc(nv, sizeof(double)), *a0 = calloc(nv, sizeof(double));
    double *d = calloc((size_t)6 * nbod * 12, sizeof(double));
    double *xscal = calloc((size_t)nbod, sizeof(double)), *vscal = calloc((size_t)nbod, sizeof(double));
    double h2[12];
    int rejected = 0;
#define D(q, k, n) d[((size_t)(n) * nbod + (k)) * 6 + (q)]
    for (int k = 1; k < nbod; k++) {
        double tmp1 = x0[3 * k] * x0[3 * k] + x0[3 * k + 1] * x0[3 * k + 1] + x0[3 * k + 2] * x0[3 * k + 2];
        double tmp2 = v0[3 * k] * v0[3 * k] + v0[3 * k + 1] * v0[3 * k + 1] + v0[3 * k + 2] * v0[3 * k + 2];
        xscal[k] = 1.0 / fmax(tmp1, 1e-300);
        vscal[k] = 1.0 / fmax(tmp2, 1e-300);
    }
    force(fc, x0, v0, a0);
    (*nforce)++;
    for (;;) {
        if (time + *h0 == time) { rejected = -1; break; }
        int done = 0;
        for (int n = 1; n <= 12 && !done; n++) {
            double h = *h0 / (double)n;
            double hby2 = .5 * h;
            h2[n - 1] = h * h;
            double h2by2 = .5 * h2[n - 1];
            for (int k = 1; k < nbod; k++)
                for (int q = 0; q < 3; q++) {
                    b[3 * k + q] = .5 * a0[3 * k + q];
                    c[3 * k + q] = 0.0;
                    xend[3 * k + q] = h2by2 * a0[3 * k + q] + h * v0[3 * k + q] + x0[3 * k + q];
                }
            for (int j = 2; j <= n; j++) {
                force(fc, xend, v0, a);
                (*nforce)++;
                double tmp0 = h * (double)j;
                for (int k = 1; k < nbod; k++)
                    for (int q = 0; q < 3; q++) {
                        b[3 * k + q] = b[3 * k + q] + a[3 * k + q];
                        c[3 * k + q] = c[3 * k + q] + b[3 * k + q];
                        xend[3 * k + q] = h2[n - 1] * c[3 * k + q] + h2by2 * a0[3 * k + q] + tmp0 * v0[3 * k + q]
                                        + x0[3 * k + q];
                    }
            }
            force(fc, xend, v0, a);
            (*nforce)++;
            for (int k = 1; k < nbod; k++)
                for (int q = 0; q < 3; q++) {
                    D(q, k, n - 1) = xend[3 * k + q];
                    D(3 + q, k, n - 1) = h * b[3 * k + q] + hby2 * a[3 * k + q] + v0[3 * k + q];
                }
            for (int j = n - 1; j >= 1; j--) {
                int j1 = j + 1;
                double tmp0 = 1.0 / (h2[j - 1] - h2[n - 1]);
                double tmp1 = tmp0 * h2[j1 - 1];
                double tmp2 = tmp0 * h2[n - 1];
                for (int k = 1; k < nbod; k++)
                    for (int q = 0; q < 6; q++) D(q, k, j - 1) = tmp1 * D(q, k, j1 - 1) - tmp2 * D(q, k, j - 1);
            }
            if (n > 3) {
                double errmax = 0.0;
                int bad = 0;
                for (int k = 1; k < nbod; k++) {
                    for (int q = 0; q < 6; q++) bad |= !isfinite(D(q, k, 0));
                    double tmp1 = fmax(fmax(D(0, k, 0) * D(0, k, 0), D(1, k, 0) * D(1, k, 0)), D(2, k, 0) * D(2, k, 0));
                    double tmp2 = fmax(fmax(D(3, k, 0) * D(3, k, 0), D(4, k, 0) * D(4, k, 0)), D(5, k, 0) * D(5, k, 0));
                    errmax = fmax(errmax, fmax(tmp1 * xscal[k], tmp2 * vscal[k]));
                }
                if (bad) { rejected = -2; done = 2; break; }
                if (errmax <= tol2) {
                    for (int k = 1; k < nbod; k++)
                        for (int q = 0; q < 3; q++) {
                            x0[3 * k + q] = D(q, k, 0);
                            v0[3 * k + q] = D(3 + q, k, 0);
                        }
                    for (int j = 2; j <= n; j++)
                        for (int k = 1; k < nbod; k++)
                            for (int q = 0; q < 3; q++) {
                                x0[3 * k + q] = x0[3 * k + q] + D(q, k, j - 1);
                                v0[3 * k + q] = v0[3 * k + q] + D(3 + q, k, j - 1);
                            }
                    *hdid = *h0;
                    if (n >= 8) *h0 = *h0 * SHRINK;
                    if (n < 7) *h0 = *h0 * GROW;
                    done = 1;
                }
            }
        }
        if (done) break;
        rejected++;
        *h0 = *h0 * .5;
    }
#undef D
    free(xend); free(b); free(c); free(a); free(a0); free(d); free(xscal); free(vscal);
    return rejected;
}

/* mdt_bs2  mercury6_2.for:3428-3617 with mfo_hkce forces (the encounter set). */
int orc_mdt_bs2(double time, double *h0, double *hdid, double tol, int nbod, const double *m, double *x0,
                double *v0, const double *rcrit, int nce, const int *ice, const int *jce, long *nforce)
{
    struct hkce_ctx c = {nbod, m, rcrit, nce, ice, jce};
    return bs2_core(time, h0, hdid, tol, nbod, x0, v0, hkce_force, &c, nforce);
}

/* mdt_bs1  mercury6_2.for:3227-3426: 8 modified-midpoint levels of 2n
 * substeps for a general (velocity-dependent) force.  Returns as bs2_core. */
static int bs1_core(double time, double *h0, double *hdid, double tol, int nbod, double *x0, double *v0,
                    bs_force_fn force, void *fc, long *nforce)
{
    const double SHRINK = .55, GROW = 1.3;
    const double tol2 = tol * tol;
    const size_t nv = (size_t)3 * nbod;
    double *x = calloc(nv, sizeof(double)), *v = calloc(nv, sizeof(double));
    double *xend = calloc(nv, sizeof(double)), *vend = calloc(nv, sizeof(double));
    double *a = calloc(nv, sizeof(double)), *a0 = calloc(nv, sizeof(double));
    double *d = calloc((size_t)6 * nbod * 8, sizeof(double));
    double *xscal = calloc((size_t)nbod, sizeof(double)), *vscal = calloc((size_t)nbod, sizeof(double));
    double h2[8];
    int rejected = 0;
#define D(q, k, n) d[((size_t)(n) * nbod + (k)) * 6 + (q)]
    for (int k = 1; k < nbod; k++) {
        double tmp1 = x0[3 * k] * x0[3 * k] + x0[3 * k + 1] * x0[3 * k + 1] + x0[3 * k + 2] * x0[3 * k + 2];
        double tmp2 = v0[3 * k] * v0[3 * k] + v0[3 * k + 1] * v0[3 * k + 1] + v0[3 * k + 2] * v0[3 * k + 2];
        xscal[k] = 1.0 / fmax(tmp1, 1e-300);
        vscal[k] = 1.0 / fmax(tmp2, 1e-300);
    }
    force(fc, x0, v0, a0);
    (*nforce)++;
    for (;;) {
        if (time + *h0 == time) { rejected = -1; break; }
        int done = 0;
        for (int n = 1; n <= 8 && !done; n++) {
            double h = *h0 / (2.0 * (double)n);
            h2[n - 1] = .25 / (double)(n * n);
            double hx2 = h * 2.0;
            for (int k = 1; k < nbod; k++)
                for (int q = 0; q < 3; q++) {
                    x[3 * k + q] = x0[3 * k + q] + h * v0[3 * k + q];
                    v[3 * k + q] = v0[3 * k + q] + h * a0[3 * k + q];
                }
            force(fc, x, v, a);
            (*nforce)++;
            for (int k = 1; k < nbod; k++)
                for (int q = 0; q < 3; q++) {
                    xend[3 * k + q] = x0[3 * k + q] + hx2 * v[3 * k + q];
                    vend[3 * k + q] = v0[3 * k + q] + hx2 * a[3 * k + q];
                }
            for (int j = 2; j <= n; j++) {
                force(fc, xend, vend, a);
                (*nforce)++;
                for (int k = 1; k < nbod; k++)
                    for (int q = 0; q < 3; q++) {
                        x[3 * k + q] = x[3 * k + q] + hx2 * vend[3 * k + q];
                        v[3 * k + q] = v[3 * k + q] + hx2 * a[3 * k + q];
                    }
                force(fc, x, v, a);
                (*nforce)++;
                for (int k = 1; k < nbod; k++)
                    for (int q = 0; q < 3; q++) {
                        xend[3 * k + q] = xend[3 * k + q] + hx2 * v[3 * k + q];
                        vend[3 * k + q] = vend[3 * k + q] + hx2 * a[3 * k + q];
                    }
            }
            force(fc, xend, vend, a);
            (*nforce)++;
            for (int k = 1; k < nbod; k++)
                for (int q = 0; q < 3; q++) {
                    D(q, k, n - 1) = .5 * (xend[3 * k + q] + x[3 * k + q] + h * vend[3 * k + q]);
                    D(3 + q, k, n - 1) = .5 * (vend[3 * k + q] + v[3 * k + q] + h * a[3 * k + q]);
                }
            for (int j = n - 1; j >= 1; j--) {
                int j1 = j + 1;
                double tmp0 = 1.0 / (h2[j - 1] - h2[n - 1]);
                double tmp1 = tmp0 * h2[j1 - 1];
                double tmp2 = tmp0 * h2[n - 1];
                for (int k = 1; k < nbod; k++)
                    for (int q = 0; q < 6; q++) D(q, k, j - 1) = tmp1 * D(q, k, j1 - 1) - tmp2 * D(q, k, j - 1);
            }
            if (n > 3) {
                double errmax = 0.0;
                int bad = 0;
                for (int k = 1; k < nbod; k++) {
                    for (int q = 0; q < 6; q++) bad |= !isfinite(D(q, k, 0));
                    double tmp1 = fmax(fmax(D(0, k, 0) * D(0, k, 0), D(1, k, 0) * D(1, k, 0)), D(2, k, 0) * D(2, k, 0));
                    double tmp2 = fmax(fmax(D(3, k, 0) * D(3, k, 0), D(4, k, 0) * D(4, k, 0)), D(5, k, 0) * D(5, k, 0));
                    errmax = fmax(errmax, fmax(tmp1 * xscal[k], tmp2 * vscal[k]));
                }
                if (bad) { rejected = -2; done = 2; break; }
                if (errmax <= tol2) {
                    for (int k = 1; k < nbod; k++)
                        for (int q = 0; q < 3; q++) {
                            x0[3 * k + q] = D(q, k, 0);
                            v0[3 * k + q] = D(3 + q, k, 0);
                        }
                    for (int j = 2; j <= n; j++)
                        for (int k = 1; k < nbod; k++)
                            for (int q = 0; q < 3; q++) {
                                x0[3 * k + q] = x0[3 * k + q] + D(q, k, j - 1);
                                v0[3 * k + q] = v0[3 * k + q] + D(3 + q, k, j - 1);
                            }
                    *hdid = *h0;
                    if (n == 8) *h0 = *h0 * SHRINK;
                    if (n < 7) *h0 = *h0 * GROW;
                    done = 1;
                }
            }
        }
        if (done) break;
        rejected++;
        *h0 = *h0 * .5;
    }
#undef D
    free(x); free(v); free(xend); free(vend); free(a); free(a0); free(d); free(xscal); free(vscal);
    return rejected;
}

/* One step of algorithm 2 (mdt_bs1) or 3 (mdt_bs2) over the whole system with
 * mfo_all (gravity, 1PN if pn, oblateness if jcen) as the driver passes it. */
int orc_mdt_bs(int algor, double time, double *h0, double *hdid, double tol, const double *jcen, int nbod,
               int nbig, const double *m, double *x0, double *v0, int pn, long *nforce)
{
    struct all_ctx c = {jcen, nbod, nbig, m, pn};
    if (algor == 2) return bs1_core(time, h0, hdid, tol, nbod, x0, v0, all_force, &c, nforce);
    return bs2_core(time, h0, hdid, tol, nbod, x0, v0, all_force, &c, nforce);
}

/* nsteps steps of algorithm 2 or 3; the sizes done in hdid[nsteps], the next
 * proposal in *hnext, halvings in *rejected.  Returns the force count, or a
 * negative bs*_core code. */
long orc_bs_run(int algor, double time0, double h0, double tol, const double *jcen, int nbod, int nbig,
                const double *m, double *x, double *v, int pn, int nsteps, double *hdid, double *hnext,
                long *rejected)
{
    long nforce = 0;
    double time = time0, h = h0;
    *rejected = 0;
    for (int s = 0; s < nsteps; s++) {
        double hd = 0.0;
        int r = orc_mdt_bs(algor, time, &h, &hd, tol, jcen, nbod, nbig, m, x, v, pn, &nforce);
        if (r < 0) return r;
        *rejected += r;
        hdid[s] = hd;
        time = time + hd;
    }
    *hnext = h;
    return nforce;
}

/* mdt_hkce  mercury6_2.for:3798-3960, collisions not allowed: the encounter
 * members are integrated by BS2 substeps over h0, mce_stat after each. */
static int hkce(double time, double h0, double *hrec, double tol, double rcen, int nbod, int nbig,
                const double *m, double *x, double *v, const double *rcrit, const double *rce,
                const double *rphys, const int *ce, int nce, const int *ice, const int *jce,
                orc_events *ev, long *nforce)
{
    int nbs = 1, nbsbig = 0;
    int *index = calloc((size_t)nbod + 1, sizeof(int)), *iback = calloc((size_t)nbod + 1, sizeof(int));
    double *mbs = calloc((size_t)nbod + 1, sizeof(double)), *xbs = calloc((size_t)3 * (nbod + 1), sizeof(double));
    double *vbs = calloc((size_t)3 * (nbod + 1), sizeof(double)), *rcb = calloc((size_t)nbod + 1, sizeof(double));
    double *rceb = calloc((size_t)nbod + 1, sizeof(double)), *rphb = calloc((size_t)nbod + 1, sizeof(double));
    double *x0 = calloc((size_t)3 * (nbod + 1), sizeof(double)), *v0 = calloc((size_t)3 * (nbod + 1), sizeof(double));
    int *ibs = calloc((size_t)(nce > 0 ? nce : 1), sizeof(int)), *jbs = calloc((size_t)(nce > 0 ? nce : 1), sizeof(int));
    int rej = 0;
    mbs[0] = m[0];
    for (int j = 1; j < nbod; j++) {
        if (ce[j] == 0) continue;
        int k = nbs++;
        if (j < nbig) nbsbig = nbs;
        mbs[k] = m[j];
        for (int q = 0; q < 3; q++) { xbs[3 * k + q] = x[3 * j + q]; vbs[3 * k + q] = v[3 * j + q]; }
        rcb[k] = rcrit[j]; rceb[k] = rce[j]; rphb[k] = rphys[j];
        index[k] = j;
        iback[j] = k;
    }
    for (int k = 0; k < nce; k++) { ibs[k] = iback[ice[k] - 1] + 1; jbs[k] = iback[jce[k] - 1] + 1; }
    int nclo0 = ev->nclo;
    double tlocal = 0.0;
    double hlocal = fsign(*hrec, h0);
    do {
        double tmp0 = fabs(h0) - fabs(tlocal);
        *hrec = hlocal;
        if (fabs(hlocal) > tmp0) hlocal = fsign(tmp0, h0);
        memcpy(x0, xbs, sizeof(double) * 3 * nbs);
        memcpy(v0, vbs, sizeof(double) * 3 * nbs);
        double hdid;
        int r = orc_mdt_bs2(tlocal, &hlocal, &hdid, tol, nbs, mbs, xbs, vbs, rcb, nce, ibs, jbs, nforce);
        if (r < 0) { rej = -1; break; }
        rej += r;
        tlocal = tlocal + hdid;
        double temp = time + tlocal;
        orc_mce_stat(temp, hdid, rcen, nbs, nbsbig, mbs, x0, v0, xbs, vbs, rceb, rphb, ev);
    } while ((tlocal - h0) * h0 < 0.0);
    for (int k = 1; k < nbs; k++) {
        int j = index[k];
        for (int q = 0; q < 3; q++) { x[3 * j + q] = xbs[3 * k + q]; v[3 * j + q] = vbs[3 * k + q]; }
    }
    for (int k = nclo0; k < ev->nclo; k++) {
        ev->iclo[k] = index[ev->iclo[k] - 1] + 1;
        ev->jclo[k] = index[ev->jclo[k] - 1] + 1;
    }
    free(index); free(iback); free(mbs); free(xbs); free(vbs); free(rcb); free(rceb); free(rphb);
    free(x0); free(v0); free(ibs); free(jbs);
    return rej;
}

/* mdt_hy  mercury6_2.for:3618-3797 (no user force; cometary forces from the
 * orc_set_ngf model: angf by mfo_ngf at a reset and after the second
 * Sun-momentum drift, 3687 and 3771, added in both kicks; the encounter
 * members' BS2 force mfo_hkce has none).
 * st->a carries the interaction accelerations; st->dtflag, st->hrec as the
 * routine's saved state.  Returns the encounter pairs found (nce). */
int orc_mdt_hy(double time, double h0, double tol, const double *jcen, double rcen, int nbod, int nbig,
               const double *m, double *x, double *v, const double *rcrit, const double *rce,
               const double *rphys, orc_hybrid *st, orc_events *ev)
{
    const double hby2 = h0 * .5;
    ev->nclo = 0;
    double *angf = (g_angf && (g_ngflag == 1 || g_ngflag == 3)) ? g_angf : NULL;
    double *zero = angf ? NULL : calloc((size_t)3 * nbod, sizeof(double));
    double *ang = angf ? angf : zero;
    if (st->dtflag != 2) {
        if (st->dtflag == 0) st->hrec = h0;
        orc_mfo_hy(jcen, nbod, nbig, m, x, rcrit, st->a);
        st->nforce++;
        st->dtflag = 2;
        if (angf) orc_mfo_ngf(nbod, x, v, angf, g_ngf);
    }
    for (int j = 1; j < nbod; j++)
        for (int q = 0; q < 3; q++) v[3 * j + q] = v[3 * j + q] + hby2 * (ang[3 * j + q] + 0.0 + st->a[3 * j + q]);
    double mv[3] = {0, 0, 0};
    for (int j = 1; j < nbod; j++)
        for (int q = 0; q < 3; q++) mv[q] = mv[q] + m[j] * v[3 * j + q];
    double temp = hby2 / m[0];
    for (int q = 0; q < 3; q++) mv[q] = temp * mv[q];
    for (int j = 1; j < nbod; j++)
        for (int q = 0; q < 3; q++) x[3 * j + q] = x[3 * j + q] + mv[q];
    double *x0 = malloc(sizeof(double) * 3 * (size_t)nbod), *v0 = malloc(sizeof(double) * 3 * (size_t)nbod);
    memcpy(x0, x, sizeof(double) * 3 * (size_t)nbod);
    memcpy(v0, v, sizeof(double) * 3 * (size_t)nbod);
    for (int j = 1; j < nbod; j++) {
        double xv[6] = {x[3 * j], x[3 * j + 1], x[3 * j + 2], v[3 * j], v[3 * j + 1], v[3 * j + 2]};
        orc_drift_one(m[0], xv, h0);
        for (int q = 0; q < 3; q++) { x[3 * j + q] = xv[q]; v[3 * j + q] = xv[3 + q]; }
    }
    int *ce = calloc((size_t)nbod, sizeof(int));
    int cap = st->cap;
    int nce = orc_mce_snif(h0, nbod, nbig, x0, v0, x, v, rcrit, ce, cap, st->ice, st->jce);
    if (nce < 0) { free(x0); free(v0); free(ce); free(zero); return -1; }
    if (nce > 0) {
        for (int j = 1; j < nbod; j++)
            if (ce[j] != 0)
                for (int q = 0; q < 3; q++) { x[3 * j + q] = x0[3 * j + q]; v[3 * j + q] = v0[3 * j + q]; }
        if (hkce(time, h0, &st->hrec, tol, rcen, nbod, nbig, m, x, v, rcrit, rce, rphys, ce, nce, st->ice,
                 st->jce, ev, &st->nforce) < 0) {
            free(x0); free(v0); free(ce); free(zero);
            return -1;
        }
    }
    mv[0] = mv[1] = mv[2] = 0.0;
    for (int j = 1; j < nbod; j++)
        for (int q = 0; q < 3; q++) mv[q] = mv[q] + m[j] * v[3 * j + q];
    for (int q = 0; q < 3; q++) mv[q] = temp * mv[q];
    for (int j = 1; j < nbod; j++)
        for (int q = 0; q < 3; q++) x[3 * j + q] = x[3 * j + q] + mv[q];
    orc_mfo_hy(jcen, nbod, nbig, m, x, rcrit, st->a);
    st->nforce++;
    if (angf) orc_mfo_ngf(nbod, x, v, angf, g_ngf);
    for (int j = 1; j < nbod; j++)
        for (int q = 0; q < 3; q++) v[3 * j + q] = v[3 * j + q] + hby2 * (ang[3 * j + q] + 0.0 + st->a[3 * j + q]);
    free(x0); free(v0); free(ce); free(zero);
    return nce;
}

/* mco_dh2h  mercury6_2.for:2007-2063: democratic heliocentric -> heliocentric */
void orc_mco_dh2h(int nbod, const double *m, const double *x, const double *v, double *xh, double *vh)
{
    double mv[3] = {0, 0, 0};
    for (int j = 1; j < nbod; j++)
        for (int q = 0; q < 3; q++) {
            xh[3 * j + q] = x[3 * j + q];
            mv[q] = mv[q] + m[j] * v[3 * j + q];
        }
    double temp = 1.0 / m[0];
    for (int q = 0; q < 3; q++) mv[q] = temp * mv[q];
    for (int j = 1; j < nbod; j++)
        for (int q = 0; q < 3; q++) vh[3 * j + q] = v[3 * j + q] + mv[q];
    for (int q = 0; q < 3; q++) { xh[q] = 0.0; vh[q] = 0.0; }
}

/* nsteps HYBRID steps as the driver takes them (mal_hcon with algorithm 10,
 * collisions not allowed): mdt_hy, then mce_cent on the physical states.
 * Event rows as orc_mvs_run (kind 1 from the encounter substeps, 8 central). */
int orc_hy_run(double time0, double h0, double tol, double rcen, const double *jcen, int nbod, int nbig,
               const double *m, double *x, double *v, const double *rcrit, const double *rce,
               const double *rphys, int nsteps, int cap, int *evi, double *evd, long *nforce)
{
    orc_hybrid st;
    st.a = calloc((size_t)3 * nbod, sizeof(double));
    st.dtflag = 0;
    st.hrec = h0;
    st.nforce = 0;
    /* encounter-pair and event capacities: at least 20000, and room for two
       pairs per body (half of a 200k-particle system near planet orbits) */
    st.cap = nbod * 2 > 20000 ? nbod * 2 : 20000;
    st.ice = calloc((size_t)st.cap, sizeof(int));
    st.jce = calloc((size_t)st.cap, sizeof(int));
    orc_events ev;
    memset(&ev, 0, sizeof(ev));
    ev.cap = st.cap;
    ev.iclo = calloc((size_t)ev.cap, sizeof(int)); ev.jclo = calloc((size_t)ev.cap, sizeof(int));
    ev.dclo = calloc((size_t)ev.cap, sizeof(double)); ev.tclo = calloc((size_t)ev.cap, sizeof(double));
    ev.ixvclo = calloc((size_t)6 * ev.cap, sizeof(double)); ev.jxvclo = calloc((size_t)6 * ev.cap, sizeof(double));
    ev.ihit = calloc((size_t)ev.cap, sizeof(int)); ev.jhit = calloc((size_t)ev.cap, sizeof(int));
    ev.chit = calloc((size_t)ev.cap, sizeof(int)); ev.dhit = calloc((size_t)ev.cap, sizeof(double));
    ev.thit = calloc((size_t)ev.cap, sizeof(double));
    double *xh0 = calloc((size_t)3 * nbod, sizeof(double)), *vh0 = calloc((size_t)3 * nbod, sizeof(double));
    double *xh = calloc((size_t)3 * nbod, sizeof(double)), *vh = calloc((size_t)3 * nbod, sizeof(double));
    int *cj = calloc((size_t)nbod, sizeof(int));
    double *ct = calloc((size_t)nbod, sizeof(double)), *cd = calloc((size_t)nbod, sizeof(double));
    double time = time0;
    int nev = 0, bad = 0;
    for (int s = 0; s < nsteps && !bad; s++) {
        orc_mco_dh2h(nbod, m, x, v, xh0, vh0);
        ev.nhit = 0;
        if (orc_mdt_hy(time, h0, tol, jcen, rcen, nbod, nbig, m, x, v, rcrit, rce, rphys, &st, &ev) < 0) {
            bad = 1;
            break;
        }
        time = time + h0;
        if (ev.overflow) bad = 1;
        for (int k = 0; k < ev.nclo; k++) {
            if (nev >= cap) { bad = 1; break; }
            evi[4 * nev] = s; evi[4 * nev + 1] = 1; evi[4 * nev + 2] = ev.iclo[k]; evi[4 * nev + 3] = ev.jclo[k];
            evd[2 * nev] = ev.dclo[k]; evd[2 * nev + 1] = ev.tclo[k];
            nev++;
        }
        orc_mco_dh2h(nbod, m, x, v, xh, vh);
        int nc = orc_mce_cent(time, h0, rcen, 2, nbod, m, xh0, vh0, xh, vh, nbod, cj, ct, cd);
        for (int k = 0; k < nc; k++) {
            if (nev >= cap) { bad = 1; break; }
            evi[4 * nev] = s; evi[4 * nev + 1] = 8; evi[4 * nev + 2] = 1; evi[4 * nev + 3] = cj[k];
            evd[2 * nev] = cd[k]; evd[2 * nev + 1] = ct[k];
            nev++;
        }
    }
    *nforce = st.nforce;
    free(st.a); free(st.ice); free(st.jce);
    free(ev.iclo); free(ev.jclo); free(ev.dclo); free(ev.tclo); free(ev.ixvclo); free(ev.jxvclo);
    free(ev.ihit); free(ev.jhit); free(ev.chit); free(ev.dhit); free(ev.thit);
    free(xh0); free(vh0); free(xh); free(vh); free(cj); free(ct); free(cd);
    return bad ? -1 : nev;
}
