// Tangent-linear and adjoint of the MVS step (massless particles + Big bodies).
//
// The forward map is mdt_mvs (mercury6_2.for:3987-4097) as the MVS kernels
// compute it.  Derivatives:
//   * Kepler drift: the Jacobian of the exact two-body flow at the converged
//     universal-variable solution.  With r0 = |x|, u = x.v, alpha =
//     2 mu/r0 - v.v and the universal anomaly s of the step,
//         x' = f x + g v,  v' = fd x + gd v,
//         f = 1 - mu G2/r0,  g = dt - mu G3,  fd = -mu G1/(r r0),  gd = 1 - mu G2/r,
//     G_n(s, alpha) = s^n c_n(alpha s^2) (Stumpff functions c_n), and s
//     solves r0 G1 + u G2 + mu G3 = dt (drift_kepu_fchk's equation).  Every
//     coefficient is a function of (r0, u, alpha, mu) through s; its partials
//     follow from dG_n/ds = G_{n-1} (dG_0/ds = -alpha G_1),
//     dG_n/dalpha|_s = (n G_{n+2} - s G_{n+1}) / 2 and the implicit-function
//     derivative ds/dq = -F_q / r.  s itself comes from the step's end state
//     through the identity s = (x'.v' - x.v + alpha dt) / mu (d(x.v)/ds =
//     mu - alpha r).
//   * Particle acceleration (mfo_mvs small-body part, jcen = 0):
//         a = a0tp - sum_p m_p K(x - X_p),  K(d) = d / |d|^3,
//         dK/dd = I/|d|^3 - 3 d d^T / |d|^5 (symmetric).
//   * Big bodies: forward mode with dual numbers through the same
//     arithmetic as the planet kernel (kick, mco_h2j, drift in Jacobi
//     coordinates with gm(m), mco_j2h, Big mfo_mvs, kick), the drift tangent
//     from the f,g Jacobian above including d/dmu.
// Per table (the planet positions, masses and indirect term a particle step
// reads) a row of R = 4 np + 3 numbers links the two sides:
//   [dX_p (3 np), dm_p (np), da0tp (3)]
// The particle adjoint sweep produces, per table, the loss sensitivity to
// exactly those numbers; the planets' tangent pass produces their
// derivatives along each parameter direction; the dot product is the
// planets' gradient.
#pragma once
#include <hip/hip_runtime.h>
#include "mvs_device.hpp"

namespace mhip {

// ---------------------------------------------------------------- Kepler Jacobian
__host__ __device__ inline void stumpff6(double z, double* c) {
    if (fabs(z) < 1.0) {
        // c4 and c5 by their series (Horner, 1/(n + 2j)! as constants), then
        // the recurrence c_k = 1/k! - z c_{k+2} downwards (no cancellation
        // for |z| < 1)
        const double f4[12] = {1.0 / 24, 1.0 / 720, 1.0 / 40320, 1.0 / 3628800, 1.0 / 479001600,
                               1.0 / 87178291200.0, 1.0 / 20922789888000.0, 1.0 / 6402373705728000.0,
                               1.0 / 2.43290200817664e18, 1.0 / 1.1240007277776077e21,
                               1.0 / 6.204484017332394e23, 1.0 / 4.0329146112660565e26};
        const double f5[12] = {1.0 / 120, 1.0 / 5040, 1.0 / 362880, 1.0 / 39916800, 1.0 / 6227020800.0,
                               1.0 / 1307674368000.0, 1.0 / 355687428096000.0, 1.0 / 1.21645100408832e17,
                               1.0 / 5.109094217170944e19, 1.0 / 2.585201673888498e22,
                               1.0 / 1.5511210043330986e25, 1.0 / 1.0888869450418352e28};
        double mz = -z;
        double c4 = f4[11], c5 = f5[11];
        for (int j = 10; j >= 0; j--) {
            c4 = f4[j] + mz * c4;
            c5 = f5[j] + mz * c5;
        }
        c[4] = c4;
        c[5] = c5;
        c[3] = 1.0 / 6.0 - z * c5;
        c[2] = 0.5 - z * c4;
        c[1] = 1.0 - z * c[3];
        c[0] = 1.0 - z * c[2];
        return;
    }
    if (z > 0.0) {
        double w = sqrt(z);
        c[0] = cos(w);
        c[1] = sin(w) / w;
    } else {
        double w = sqrt(-z);
        c[0] = cosh(w);
        c[1] = sinh(w) / w;
    }
    double iz = 1.0 / z;
    c[2] = (1.0 - c[0]) * iz;
    c[3] = (1.0 - c[1]) * iz;
    c[4] = (0.5 - c[2]) * iz;
    c[5] = (1.0 / 6.0 - c[3]) * iz;
}

// Lagrange coefficients and their partials with respect to q = (r0, u, alpha, mu).
struct KepSens {
    double f, g, fd, gd;
    double fq[4], gq[4], fdq[4], gdq[4];
    double r0, alpha;
};

enum { Q_R0 = 0, Q_U = 1, Q_AL = 2, Q_MU = 3 };

__host__ __device__ inline void kepler_sens(double mu, double dt, const double* x, const double* v, double s,
                                            KepSens& K) {
    double r0 = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    double ir0 = 1.0 / r0;
    double u = x[0] * v[0] + x[1] * v[1] + x[2] * v[2];
    double alpha = 2.0 * mu * ir0 - (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double c[6];
    stumpff6(alpha * s * s, c);
    double G[6];
    double sp = 1.0;
    for (int n = 0; n < 6; n++) { G[n] = sp * c[n]; sp = sp * s; }
    double Ga[4];
    for (int n = 0; n < 4; n++) Ga[n] = (n * G[n + 2] - s * G[n + 1]) * 0.5;
    double r = r0 * G[0] + u * G[1] + mu * G[2];
    double ir = 1.0 / r;
    double Fq[4] = {G[1], G[2], r0 * Ga[1] + u * Ga[2] + mu * Ga[3], G[3]};
    double sq[4];
    for (int q = 0; q < 4; q++) sq[q] = -Fq[q] * ir;
    // total derivatives of G_0..G_3
    double dG[4][4];
    for (int n = 0; n < 4; n++)
        for (int q = 0; q < 4; q++) {
            double ds = (n >= 1 ? G[n - 1] : -alpha * G[1]) * sq[q];
            dG[n][q] = ds + (q == Q_AL ? Ga[n] : 0.0);
        }
    double rq[4];
    for (int q = 0; q < 4; q++)
        rq[q] = r0 * dG[0][q] + u * dG[1][q] + mu * dG[2][q] + (q == Q_R0 ? G[0] : 0.0) + (q == Q_U ? G[1] : 0.0)
              + (q == Q_MU ? G[2] : 0.0);
    const double irr0 = ir * ir0;
    K.f = 1.0 - mu * G[2] * ir0;
    K.g = dt - mu * G[3];
    K.fd = -mu * G[1] * irr0;
    K.gd = 1.0 - mu * G[2] * ir;
    for (int q = 0; q < 4; q++) {
        K.fq[q] = -(mu * ir0) * dG[2][q] + (q == Q_R0 ? mu * G[2] * ir0 * ir0 : 0.0) + (q == Q_MU ? -G[2] * ir0 : 0.0);
        K.gq[q] = -mu * dG[3][q] + (q == Q_MU ? -G[3] : 0.0);
        K.fdq[q] = -mu * dG[1][q] * irr0 + mu * G[1] * (rq[q] * r0 + (q == Q_R0 ? r : 0.0)) * irr0 * irr0
                 + (q == Q_MU ? -G[1] * irr0 : 0.0);
        K.gdq[q] = -mu * dG[2][q] * ir + mu * G[2] * rq[q] * ir * ir + (q == Q_MU ? -G[2] * ir : 0.0);
    }
    K.r0 = r0;
    K.alpha = alpha;
}

// Universal anomaly of a drift from (x, v) to (x1, v1) over dt.
__host__ __device__ inline double kepler_anomaly(double mu, double dt, const double* x, const double* v,
                                                 const double* x1, const double* v1) {
    double r0 = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    double alpha = 2.0 * mu / r0 - (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double u0 = x[0] * v[0] + x[1] * v[1] + x[2] * v[2];
    double u1 = x1[0] * v1[0] + x1[1] * v1[1] + x1[2] * v1[2];
    return (u1 - u0 + alpha * dt) / mu;
}

// Tangent of the drift: (dx, dv, dmu) -> (dx1, dv1).
__host__ __device__ inline void kepler_tangent(const KepSens& K, double mu, const double* x, const double* v,
                                               const double* dx, const double* dv, double dmu, double* dx1,
                                               double* dv1) {
    double r0 = K.r0;
    double dr0 = (x[0] * dx[0] + x[1] * dx[1] + x[2] * dx[2]) / r0;
    double du = dx[0] * v[0] + dx[1] * v[1] + dx[2] * v[2] + x[0] * dv[0] + x[1] * dv[1] + x[2] * dv[2];
    double dal = 2.0 * dmu / r0 - 2.0 * mu * dr0 / (r0 * r0) - 2.0 * (v[0] * dv[0] + v[1] * dv[1] + v[2] * dv[2]);
    double dq[4] = {dr0, du, dal, dmu};
    double df = 0, dg = 0, dfd = 0, dgd = 0;
    for (int q = 0; q < 4; q++) {
        df += K.fq[q] * dq[q];
        dg += K.gq[q] * dq[q];
        dfd += K.fdq[q] * dq[q];
        dgd += K.gdq[q] * dq[q];
    }
    for (int c = 0; c < 3; c++) {
        dx1[c] = K.f * dx[c] + K.g * dv[c] + df * x[c] + dg * v[c];
        dv1[c] = K.fd * dx[c] + K.gd * dv[c] + dfd * x[c] + dgd * v[c];
    }
}

// Adjoint of the drift: (lx1, lv1) -> (lx, lv) at fixed mu.
__host__ __device__ inline void kepler_adjoint(const KepSens& K, double mu, const double* x, const double* v,
                                               const double* lx1, const double* lv1, double* lx, double* lv) {
    double A = x[0] * lx1[0] + x[1] * lx1[1] + x[2] * lx1[2];
    double B = v[0] * lx1[0] + v[1] * lx1[1] + v[2] * lx1[2];
    double C = x[0] * lv1[0] + x[1] * lv1[1] + x[2] * lv1[2];
    double D = v[0] * lv1[0] + v[1] * lv1[1] + v[2] * lv1[2];
    double sg[3];
    for (int q = 0; q < 3; q++) sg[q] = A * K.fq[q] + B * K.gq[q] + C * K.fdq[q] + D * K.gdq[q];
    double r0 = K.r0;
    double cx = sg[Q_R0] / r0 - 2.0 * mu * sg[Q_AL] / (r0 * r0 * r0);
    for (int c = 0; c < 3; c++) {
        lx[c] = K.f * lx1[c] + K.fd * lv1[c] + cx * x[c] + sg[Q_U] * v[c];
        lv[c] = K.g * lx1[c] + K.gd * lv1[c] + sg[Q_U] * x[c] - 2.0 * sg[Q_AL] * v[c];
    }
}

// ---------------------------------------------------------------- dual numbers
struct Dual {
    double v, d;
};
__device__ __forceinline__ Dual dual(double v, double d = 0.0) { return Dual{v, d}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual{a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual{a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    double q = a.v / b.v;
    return Dual{q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator/(double a, Dual b) {
    double q = a / b.v;
    return Dual{q, -q * b.d / b.v};
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual{a.v + b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual{a - b.v, -b.d}; }
__device__ __forceinline__ Dual dsqrt(Dual a) {
    double s = sqrt(a.v);
    return Dual{s, a.d * 0.5 / s};
}

// ---------------------------------------------------------------- planets, forward mode
// Row of one table: [dX_p (3 np), dm_p (np), da0tp (3)].
__host__ __device__ inline int row_len(int np) { return 4 * np + 3; }

struct TlmArgs {
    const Sys* sys;
    const double* m;    // masses (body-indexed, system base included)
    const double* px;   // initial planet positions [3][np]
    const double* pv;   // initial planet velocities [3][np]
    int steps;
    const double* seed; // one direction [3np dX, 3np dV, np dm]; null: unit direction blockIdx.x
    double* rows;       // per-table rows [ndir][(steps + 1)][R], start table first
    double* fin;        // final planet tangents [ndir][3np dX, 3np dV] (may be null)
};

// Big-body accelerations of mfo_mvs (the planet kernel's big_forces, jcen = 0)
// in dual arithmetic; lane p owns planet p.  Also the indirect term a0tp.
__device__ inline void big_forces_dual(int np, const Dual* m, const Dual* minside, const Dual* X, const Dual* XJ,
                                       Dual* A, Dual* t0, Dual* t2, Dual* a1, Dual* a0tp) {
    const int p = threadIdx.x;
    if (p < np) {
        if (p >= 1) {
            const Dual* xk = X + 3 * p;
            const Dual* xjk = XJ + 3 * p;
            Dual r2 = xk[0] * xk[0] + xk[1] * xk[1] + xk[2] * xk[2];
            Dual rj2 = xjk[0] * xjk[0] + xjk[1] * xjk[1] + xjk[2] * xjk[2];
            Dual r = 1.0 / dsqrt(r2);
            Dual rj = 1.0 / dsqrt(rj2);
            Dual r3 = r * r * r;
            Dual rj3 = rj * rj * rj;
            Dual mk = m[p + 1];
            Dual fac0 = mk * r3;
            Dual fac12 = m[0] * rj3;
            Dual fac2 = mk * fac12 / (minside[p] + m[0]);
            Dual q = (r2 - rj2) * .5 / rj2;
            Dual q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q3 * q4;
            Dual fac1 = 402.1875 * q7 - 187.6875 * q6 + 86.625 * q5 - 39.375 * q4 + 17.5 * q3 - 7.5 * q2
                      + 3.0 * q - dual(1.0);
            for (int c = 0; c < 3; c++) {
                t0[3 * p + c] = fac0 * xk[c];
                a1[3 * p + c] = fac12 * (xjk[c] + fac1 * xk[c]);
                t2[3 * p + c] = fac2 * xjk[c];
            }
        } else {
            for (int c = 0; c < 3; c++) { t0[c] = dual(0); a1[c] = dual(0); t2[c] = dual(0); }
        }
        Dual a3[3] = {dual(0), dual(0), dual(0)};
        const Dual* xb = X + 3 * p;
        for (int i = 0; i < p; i++) {
            Dual dx = xb[0] - X[3 * i], dy = xb[1] - X[3 * i + 1], dz = xb[2] - X[3 * i + 2];
            Dual s_1 = 1.0 / dsqrt(dx * dx + dy * dy + dz * dz);
            Dual faci = m[i + 1] * (s_1 * s_1 * s_1);
            a3[0] = a3[0] - faci * dx; a3[1] = a3[1] - faci * dy; a3[2] = a3[2] - faci * dz;
        }
        for (int j = p + 1; j < np; j++) {
            Dual dx = X[3 * j] - xb[0], dy = X[3 * j + 1] - xb[1], dz = X[3 * j + 2] - xb[2];
            Dual s_1 = 1.0 / dsqrt(dx * dx + dy * dy + dz * dz);
            Dual facj = m[j + 1] * (s_1 * s_1 * s_1);
            a3[0] = a3[0] + facj * dx; a3[1] = a3[1] + facj * dy; a3[2] = a3[2] + facj * dz;
        }
        for (int c = 0; c < 3; c++) A[3 * p + c] = a3[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Dual a0[3] = {dual(0), dual(0), dual(0)}, a2[3] = {dual(0), dual(0), dual(0)};
        for (int q = 1; q < np; q++)
            for (int c = 0; c < 3; c++) a0[c] = a0[c] - t0[3 * q + c];
        for (int c = 0; c < 3; c++) a0tp[c] = dual(0);
        if (np >= 1) {
            Dual r = 1.0 / dsqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
            Dual fac0 = m[1] * (r * r * r);
            for (int c = 0; c < 3; c++) a0tp[c] = a0[c] - fac0 * X[c];
        }
        for (int q = 0; q < np; q++)
            for (int c = 0; c < 3; c++) {
                if (q >= 1) a2[c] = a2[c] + t2[3 * q + c];
                A[3 * q + c] = a0[c] + a1[3 * q + c] + (q >= 1 ? a2[c] : dual(0)) + A[3 * q + c];
            }
    }
    __syncthreads();
}

__device__ inline void h2j_dual(int np, const Dual* m, const Dual* jdx, const Dual* X, const Dual* V, Dual* XJ,
                                Dual* VJ) {
    if (np < 1) return;
    for (int k = 0; k < 3; k++) { XJ[k] = X[k]; VJ[k] = V[k]; }
    Dual mx[3], mu[3];
    for (int k = 0; k < 3; k++) { mx[k] = m[1] * X[k]; mu[k] = m[1] * V[k]; }
    for (int p = 1; p < np; p++) {
        for (int k = 0; k < 3; k++) {
            XJ[3 * p + k] = X[3 * p + k] - jdx[p] * mx[k];
            VJ[3 * p + k] = V[3 * p + k] - jdx[p] * mu[k];
        }
        if (p < np - 1)
            for (int k = 0; k < 3; k++) {
                mx[k] = mx[k] + m[p + 1] * X[3 * p + k];
                mu[k] = mu[k] + m[p + 1] * V[3 * p + k];
            }
    }
}

__device__ inline void j2h_dual(int np, const Dual* jhx, const Dual* XJ, const Dual* VJ, Dual* X, Dual* V) {
    if (np < 1) return;
    for (int k = 0; k < 3; k++) { X[k] = XJ[k]; V[k] = VJ[k]; }
    Dual mx[3], mu[3];
    for (int k = 0; k < 3; k++) { mx[k] = jhx[0] * XJ[k]; mu[k] = jhx[0] * VJ[k]; }
    for (int p = 1; p < np; p++) {
        for (int k = 0; k < 3; k++) {
            X[3 * p + k] = XJ[3 * p + k] + mx[k];
            V[3 * p + k] = VJ[3 * p + k] + mu[k];
        }
        if (p < np - 1)
            for (int k = 0; k < 3; k++) {
                mx[k] = mx[k] + jhx[p] * XJ[3 * p + k];
                mu[k] = mu[k] + jhx[p] * VJ[3 * p + k];
            }
    }
}

// Write the tangent row of one table (lane p: its planet; lane 0: a0tp).
__device__ __forceinline__ void table_row(int np, const Dual* X, const Dual* m, const Dual* a0tp, double* row) {
    const int p = threadIdx.x;
    if (p < np) {
        for (int c = 0; c < 3; c++) row[3 * p + c] = X[3 * p + c].d;
        row[3 * np + p] = m[p + 1].d;
    }
    if (p == 0)
        for (int c = 0; c < 3; c++) row[4 * np + c] = a0tp[c].d;
}

// The Big bodies' MVS steps in forward mode, one workgroup (one wave) per
// direction.  Values repeat the planet kernel's arithmetic; tangents ride
// along.  Direction d of the unit set: d < 3np position (planet d/3, comp
// d%3), then 3np velocities, then np masses.
__global__ void __launch_bounds__(64) tlm_planets(TlmArgs args) {
    const Sys d = args.sys[0];
    const int np = d.nbig - 1;
    const int lane = threadIdx.x;
    const int dir = blockIdx.x;
    const int R = row_len(np);
    double* rows = args.rows + (long long)dir * (args.steps + 1) * R;
    __shared__ Dual m[kMaxBig], minside[kMaxBig], jdx[kMaxBig], jhx[kMaxBig], gm[kMaxBig];
    __shared__ Dual X[3 * kMaxBig], V[3 * kMaxBig], A[3 * kMaxBig], XJ[3 * kMaxBig], VJ[3 * kMaxBig];
    __shared__ Dual t0[3 * kMaxBig], t2[3 * kMaxBig], a1[3 * kMaxBig], a0tp[3];
    auto seed_of = [&](int k) -> double {
        if (args.seed) return args.seed[k];
        return k == dir ? 1.0 : 0.0;
    };
    const int b0 = d.base;
    for (int q = lane; q < d.nbig; q += 64) m[q] = dual(args.m[b0 + q], q >= 1 ? seed_of(6 * np + q - 1) : 0.0);
    for (int p = lane; p < np; p += 64)
        for (int c = 0; c < 3; c++) {
            X[3 * p + c] = dual(args.px[c * np + p], seed_of(3 * p + c));
            V[3 * p + c] = dual(args.pv[c * np + p], seed_of(3 * np + 3 * p + c));
        }
    __syncthreads();
    if (lane == 0) {
        Dual ms = dual(0);
        for (int p = 0; p < np; p++) { minside[p] = ms; ms = ms + m[p + 1]; }
        Dual mtot = np >= 1 ? m[1] : dual(0);
        for (int p = 1; p < np; p++) { jdx[p] = 1.0 / (mtot + m[0]); mtot = mtot + m[p + 1]; }
        mtot = np >= 1 ? m[1] : dual(0);
        if (np >= 1) jhx[0] = m[1] / (mtot + m[0]);
        for (int p = 1; p < np; p++) { mtot = mtot + m[p + 1]; jhx[p] = m[p + 1] / (mtot + m[0]); }
        Dual mins = m[0];
        for (int p = 0; p < np; p++) {
            Dual msofar = mins + m[p + 1];
            gm[p] = m[0] * msofar / mins;
            mins = msofar;
        }
    }
    __syncthreads();
    // accelerations at the start (the planet kernel's reset path) and the start table
    if (lane == 0) h2j_dual(np, m, jdx, X, V, XJ, VJ);
    __syncthreads();
    big_forces_dual(np, m, minside, X, XJ, A, t0, t2, a1, a0tp);
    table_row(np, X, m, a0tp, rows);
    const double h = d.h, hby2 = d.hby2;
    for (int s = 0; s < args.steps; s++) {
        const int p = lane;
        if (p < np)
            for (int c = 0; c < 3; c++) V[3 * p + c] = V[3 * p + c] + hby2 * A[3 * p + c];
        __syncthreads();
        if (lane == 0) h2j_dual(np, m, jdx, X, V, XJ, VJ);
        __syncthreads();
        if (p < np) {
            // value: the planet kernel's drift; tangent: the f,g Jacobian
            Orbit o{XJ[3 * p].v, XJ[3 * p + 1].v, XJ[3 * p + 2].v, VJ[3 * p].v, VJ[3 * p + 1].v, VJ[3 * p + 2].v};
            double x0[3] = {o.x, o.y, o.z}, v0[3] = {o.vx, o.vy, o.vz};
            drift_one(gm[p].v, o, h);
            double x1[3] = {o.x, o.y, o.z}, v1[3] = {o.vx, o.vy, o.vz};
            double sanom = kepler_anomaly(gm[p].v, h, x0, v0, x1, v1);
            KepSens K;
            kepler_sens(gm[p].v, h, x0, v0, sanom, K);
            double dx[3], dv[3], dx1[3], dv1[3];
            for (int c = 0; c < 3; c++) { dx[c] = XJ[3 * p + c].d; dv[c] = VJ[3 * p + c].d; }
            kepler_tangent(K, gm[p].v, x0, v0, dx, dv, gm[p].d, dx1, dv1);
            for (int c = 0; c < 3; c++) {
                XJ[3 * p + c] = dual(x1[c], dx1[c]);
                VJ[3 * p + c] = dual(v1[c], dv1[c]);
            }
        }
        __syncthreads();
        if (lane == 0) j2h_dual(np, jhx, XJ, VJ, X, V);
        __syncthreads();
        big_forces_dual(np, m, minside, X, XJ, A, t0, t2, a1, a0tp);
        table_row(np, X, m, a0tp, rows + (long long)(s + 1) * R);
        if (p < np)
            for (int c = 0; c < 3; c++) V[3 * p + c] = V[3 * p + c] + hby2 * A[3 * p + c];
        __syncthreads();
    }
    if (args.fin) {
        double* fin = args.fin + (long long)dir * 6 * np;
        for (int p = lane; p < np; p += 64)
            for (int c = 0; c < 3; c++) {
                fin[3 * p + c] = X[3 * p + c].d;
                fin[3 * np + 3 * p + c] = V[3 * p + c].d;
            }
    }
}

// ---------------------------------------------------------------- particles
struct AdjArgs {
    const Sys* sys;
    State st;            // particle initial state (rows nbig..n-1)
    const double* tab;   // tables: step t at t * ts, the start table at steps * ts
    int steps;           // total steps of the tape
    int k0, k1;          // this launch covers steps [k0, k1)
    double* ck_in;       // particle (x, v) at step k0: [6][npart] (null: the state)
    double* ck_out;      // particle (x, v) at step k1: [6][npart] (may be null)
    double* rec;         // record: per step k0..k1-1 [7][npart]: x_k, v after the opening
                         // kick (the reverse overwrites it with the kick weight l_k), and
                         // the drift's universal anomaly s_k
    const double* xfin;  // final particle positions [3][npart] (reverse of the last step)
    double* fin_ell;     // (x_T, l_T) of the last step's closing kick [6][npart]
    double* lam;         // adjoint: (lambda_x, lambda_v) [6][npart], in/out
    double* part;        // per-block sensitivity rows [(steps + 1)][nblk][R]
    const double* rows;  // tangent: planet rows [(steps + 1)][R]
    double* tan;         // tangent: (dx, dv) [6][npart], in/out
    int npart;
};

__device__ __forceinline__ const double* table_at(const AdjArgs& a, const Sys& d, int t) {
    return a.tab + (long long)(t >= 0 ? t : a.steps) * d.ts;
}

__device__ inline void accel_of(const Sys& d, const double* T, int np, const double* x, double* a) {
    particle_accel(d, T, np, x[0], x[1], x[2], a[0], a[1], a[2]);
}

// Forward particle steps [k0, k1) from the tables; optional record of every
// step and checkpoint of the end state.  Same arithmetic as the MVS particle
// kernels (drift_one is the full drift_dan chain).
__global__ void __launch_bounds__(128) adj_forward(AdjArgs args) {
    const Sys d = args.sys[0];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= args.npart) return;
    const long long cap = args.st.cap;
    const int g = d.base + d.nbig + i;
    const int np = d.nbig - 1;
    const int n = args.npart;
    double x[3], v[3];
    for (int c = 0; c < 3; c++) {
        x[c] = args.ck_in ? args.ck_in[c * n + i] : args.st.x[c * cap + g];
        v[c] = args.ck_in ? args.ck_in[(3 + c) * n + i] : args.st.v[c * cap + g];
    }
    double a[3];
    accel_of(d, table_at(args, d, args.k0 - 1), np, x, a);
    for (int k = args.k0; k < args.k1; k++) {
        double vh[3] = {v[0] + d.hby2 * a[0], v[1] + d.hby2 * a[1], v[2] + d.hby2 * a[2]};
        Orbit o{x[0], x[1], x[2], vh[0], vh[1], vh[2]};
        drift_one(d.mu, o, d.h);
        double x1[3] = {o.x, o.y, o.z}, v1[3] = {o.vx, o.vy, o.vz};
        if (args.rec) {
            double* R = args.rec + (long long)(k - args.k0) * 7 * n;
            for (int c = 0; c < 3; c++) { R[c * n + i] = x[c]; R[(3 + c) * n + i] = vh[c]; }
            R[6 * n + i] = kepler_anomaly(d.mu, d.h, x, vh, x1, v1);
        }
        for (int c = 0; c < 3; c++) x[c] = x1[c];
        accel_of(d, table_at(args, d, k), np, x, a);
        for (int c = 0; c < 3; c++) v[c] = v1[c] + d.hby2 * a[c];
    }
    if (args.ck_out)
        for (int c = 0; c < 3; c++) { args.ck_out[c * n + i] = x[c]; args.ck_out[(3 + c) * n + i] = v[c]; }
}

// lx += (da/dx)^T l at position x with table T: da/dx = -sum_p m_p dK/dd.
__device__ inline void kick_jac(const double* __restrict__ T, int np, const double* x, const double* l, double* lx) {
    const double* F = T + kHdr;
    for (int p = 0; p < np; p++) {
        double dd[3] = {x[0] - F[kF * p], x[1] - F[kF * p + 1], x[2] - F[kF * p + 2]};
        double ir = 1.0 / sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
        double ir3 = ir * ir * ir;
        double dl = dd[0] * l[0] + dd[1] * l[1] + dd[2] * l[2];
        double mp = F[kF * p + 3];
        double k5 = 3.0 * dl * ir3 * ir * ir;
        for (int c = 0; c < 3; c++) lx[c] -= mp * (l[c] * ir3 - dd[c] * k5);
    }
}

// Reverse sweep over steps [k0, k1) (k1 - 1 first) from the recorded steps;
// lambda in registers.  Each step leaves its kick weight l_k in the record
// for adj_rows.
__global__ void __launch_bounds__(128) adj_reverse(AdjArgs args) {
    const Sys d = args.sys[0];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= args.npart) return;
    const int np = d.nbig - 1;
    const int n = args.npart;
    double lx[3], lv[3];
    for (int c = 0; c < 3; c++) { lx[c] = args.lam[c * n + i]; lv[c] = args.lam[(3 + c) * n + i]; }
    for (int k = args.k1 - 1; k >= args.k0; k--) {
        double* Rr = args.rec + (long long)(k - args.k0) * 7 * n;
        double x[3], vh[3];
        for (int c = 0; c < 3; c++) { x[c] = Rr[c * n + i]; vh[c] = Rr[(3 + c) * n + i]; }
        double s = Rr[6 * n + i];
        if (k == args.steps - 1) {
            // the last step's closing half kick at the final position
            double xT[3] = {args.xfin[i], args.xfin[n + i], args.xfin[2 * n + i]};
            double l[3] = {d.hby2 * lv[0], d.hby2 * lv[1], d.hby2 * lv[2]};
            kick_jac(table_at(args, d, k), np, xT, l, lx);
            for (int c = 0; c < 3; c++) { args.fin_ell[c * n + i] = xT[c]; args.fin_ell[(3 + c) * n + i] = l[c]; }
        }
        KepSens K;
        kepler_sens(d.mu, d.h, x, vh, s, K);
        double nx[3], nv[3];
        kepler_adjoint(K, d.mu, x, vh, lx, lv, nx, nv);
        // the kick at x_k: this step's opening half kick and the previous
        // step's closing half kick (same point, same table, same lambda_v)
        double w = k == 0 ? d.hby2 : d.h;
        double l[3] = {w * nv[0], w * nv[1], w * nv[2]};
        for (int c = 0; c < 3; c++) { lx[c] = nx[c]; lv[c] = nv[c]; }
        kick_jac(table_at(args, d, k - 1), np, x, l, lx);
        for (int c = 0; c < 3; c++) Rr[(3 + c) * n + i] = l[c];
    }
    for (int c = 0; c < 3; c++) { args.lam[c * n + i] = lx[c]; args.lam[(3 + c) * n + i] = lv[c]; }
}

constexpr int kAdjRmax = 4 * 12 + 3;   // up to 12 planets in the adjoint

// Sensitivity rows of the tables read by steps [k0, k1) (plus the final
// table when the range ends the tape): per workgroup of 128 particles,
//   dX_p += m_p (l/|d|^3 - 3 d (d.l)/|d|^5),  dm_p += -(d.l)/|d|^3,  da0tp += l
// summed over the workgroup in a fixed order.  blockIdx.x: step, y: particles.
__global__ void __launch_bounds__(128) adj_rows(AdjArgs args) {
    const Sys d = args.sys[0];
    const int np = d.nbig - 1;
    const int R = row_len(np);
    const int n = args.npart;
    const int r = blockIdx.x;
    const int i = blockIdx.y * blockDim.x + threadIdx.x;
    const int nblk = gridDim.y;
    const bool last = args.k0 + r == args.k1;      // the final table's closing kick
    const int k = last ? args.steps : args.k0 + r;  // sensitivity row index (table k - 1)
    __shared__ double lds[128 * kAdjRmax];
    double* my = lds + threadIdx.x * R;
    if (i < n) {
        const double* P = last ? args.fin_ell : args.rec + (long long)r * 7 * n;
        double x[3] = {P[i], P[n + i], P[2 * n + i]};
        double l[3] = {P[3 * n + i], P[4 * n + i], P[5 * n + i]};
        const double* F = table_at(args, d, k - 1) + kHdr;
        for (int p = 0; p < np; p++) {
            double dd[3] = {x[0] - F[kF * p], x[1] - F[kF * p + 1], x[2] - F[kF * p + 2]};
            double ir = 1.0 / sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
            double ir3 = ir * ir * ir;
            double dl = dd[0] * l[0] + dd[1] * l[1] + dd[2] * l[2];
            double mp = F[kF * p + 3];
            double k5 = 3.0 * dl * ir3 * ir * ir;
            for (int c = 0; c < 3; c++) my[3 * p + c] = mp * (l[c] * ir3 - dd[c] * k5);
            my[3 * np + p] = -dl * ir3;
        }
        for (int c = 0; c < 3; c++) my[4 * np + c] = l[c];
    } else {
        for (int q = 0; q < R; q++) my[q] = 0.0;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < R; q += blockDim.x) {
        double sum = 0.0;
        for (int t = 0; t < (int)blockDim.x; t++) sum += lds[t * R + q];
        args.part[((long long)k * nblk + blockIdx.y) * R + q] = sum;
    }
}

// src[t][q] = sum over workgroups (in order) of part[t][b][q].
__global__ void adj_sum_rows(const double* part, double* src, int ntab, int nblk, int R) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)ntab * R) return;
    const long long t = idx / R, q = idx % R;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[(t * nblk + b) * R + q];
    src[idx] = s;
}

// grad[dir] = sum over tables and row entries of src * rows[dir], one
// workgroup per direction, fixed summation order.
__global__ void __launch_bounds__(256) adj_contract(const double* src, const double* rows, double* grad, int ntab,
                                                    int R) {
    const int dir = blockIdx.x;
    const double* rw = rows + (long long)dir * ntab * R;
    const long long tot = (long long)ntab * R;
    double acc = 0.0;
    for (long long idx = threadIdx.x; idx < tot; idx += blockDim.x) acc += src[idx] * rw[idx];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[dir] = red[0];
}

// Tangent of the particle steps [k0, k1): (dx, dv) forward with the planet rows.
__device__ inline void kick_tangent(const double* __restrict__ T, const double* row, int np, const double* x,
                                    const double* dx, double w, double* dv) {
    const double* F = T + kHdr;
    double da[3] = {row[4 * np], row[4 * np + 1], row[4 * np + 2]};
    for (int p = 0; p < np; p++) {
        double dd[3] = {x[0] - F[kF * p], x[1] - F[kF * p + 1], x[2] - F[kF * p + 2]};
        double ddd[3] = {dx[0] - row[3 * p], dx[1] - row[3 * p + 1], dx[2] - row[3 * p + 2]};
        double r2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
        double ir = 1.0 / sqrt(r2);
        double ir3 = ir * ir * ir;
        double dq = dd[0] * ddd[0] + dd[1] * ddd[1] + dd[2] * ddd[2];
        double mp = F[kF * p + 3];
        double dm = row[3 * np + p];
        for (int c = 0; c < 3; c++)
            da[c] -= mp * (ddd[c] * ir3 - 3.0 * dd[c] * dq * ir3 * ir * ir) + dm * dd[c] * ir3;
    }
    for (int c = 0; c < 3; c++) dv[c] += w * da[c];
}

__global__ void __launch_bounds__(128) tlm_particles(AdjArgs args) {
    const Sys d = args.sys[0];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= args.npart) return;
    const long long cap = args.st.cap;
    const int g = d.base + d.nbig + i;
    const int np = d.nbig - 1;
    const int R = row_len(np);
    const int n = args.npart;
    double x[3], v[3], dx[3], dv[3];
    for (int c = 0; c < 3; c++) {
        x[c] = args.ck_in ? args.ck_in[c * n + i] : args.st.x[c * cap + g];
        v[c] = args.ck_in ? args.ck_in[(3 + c) * n + i] : args.st.v[c * cap + g];
        dx[c] = args.tan[c * n + i];
        dv[c] = args.tan[(3 + c) * n + i];
    }
    double a[3];
    accel_of(d, table_at(args, d, args.k0 - 1), np, x, a);
    for (int k = args.k0; k < args.k1; k++) {
        const double* rp = args.rows + (long long)k * R;          // table k - 1 (start table at row 0)
        const double* rk = args.rows + (long long)(k + 1) * R;    // table k
        kick_tangent(table_at(args, d, k - 1), rp, np, x, dx, d.hby2, dv);
        double vh[3] = {v[0] + d.hby2 * a[0], v[1] + d.hby2 * a[1], v[2] + d.hby2 * a[2]};
        Orbit o{x[0], x[1], x[2], vh[0], vh[1], vh[2]};
        drift_one(d.mu, o, d.h);
        double x1[3] = {o.x, o.y, o.z}, v1[3] = {o.vx, o.vy, o.vz};
        double s = kepler_anomaly(d.mu, d.h, x, vh, x1, v1);
        KepSens K;
        kepler_sens(d.mu, d.h, x, vh, s, K);
        double dx1[3], dv1[3];
        kepler_tangent(K, d.mu, x, vh, dx, dv, 0.0, dx1, dv1);
        for (int c = 0; c < 3; c++) { x[c] = x1[c]; dx[c] = dx1[c]; dv[c] = dv1[c]; }
        accel_of(d, table_at(args, d, k), np, x, a);
        for (int c = 0; c < 3; c++) v[c] = v1[c] + d.hby2 * a[c];
        kick_tangent(table_at(args, d, k), rk, np, x, dx, d.hby2, dv);
    }
    for (int c = 0; c < 3; c++) {
        args.tan[c * n + i] = dx[c];
        args.tan[(3 + c) * n + i] = dv[c];
        if (args.ck_out) { args.ck_out[c * n + i] = x[c]; args.ck_out[(3 + c) * n + i] = v[c]; }
    }
}

}  // namespace mhip
