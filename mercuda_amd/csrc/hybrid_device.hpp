// HYBRID (algorithm 10) for massless particles on gfx950, FP64.
//
// One regular step of mdt_hy (mercury6_2.for:3618-3797) splits, as the MVS
// step does, into a Big-body wave and a particle pass, because massless
// particles never act on Big bodies:
//   hy_big_begin      one wave: Big bodies' interaction kick (mfo_hy terms),
//                     Sun-momentum drift, Kepler drift, x-y boxes padded by
//                     rcrit, Big-Big encounter sniff; tabulates what the
//                     particles need (momentum shifts, Big start/pre-drift/
//                     post-drift states, boxes, rcrit, masses);
//   hy_particles_begin one lane per particle: kick, shift, drift, sniff
//                     against the Big boxes (mce_snif);
//   hy_big_finish / hy_particles_finish: second Sun-momentum drift, mfo_hy
//                     accelerations, kick, physical velocities.
// With the cometary model (ngflag 1 or 3) every body carries mdt_hy's angf
// in st.angf: mfo_ngf at a reset (mercury6_2.for:3687) and after the second
// Sun-momentum drift (3771), added in both kicks as hby2 * (angf + ausr + a)
// with no user force; the encounter members' mfo_hkce has no such term.
// Encounter pairs are appended by atomics and sorted into the Fortran loop
// order on the host.  The encounter members' Keplerian part (mdt_hkce) is the
// driver's, integrated through the compact context below (mdt_bs2 with
// mfo_hkce forces, mercury6_2.for:3428-3617).
#pragma once
#include <hip/hip_runtime.h>
#include "mvs_device.hpp"

namespace mhip {

// ---------------------------------------------------------------- regular map
// Per Big body p of the begin table
enum { HB_XS = 0, HB_X0 = 3, HB_V0 = 6, HB_X1 = 9, HB_V1 = 12, HB_BOX = 15, HB_RC = 19, HB_M = 20, HB_N = 21 };
// header: mv0 (start physical shift), mvs (solar shift), acen (oblateness reaction of the start state)
constexpr int kHyHdr = 12;

struct HyArgs {
    const Sys* sys;
    State st;
    const double* rcrit;
    double* x0;              // pre-drift positions [3][cap]
    double* v0;              // pre-drift velocities [3][cap]
    double* tab;             // kHyHdr + HB_N * np doubles
    int* pairs;              // encounter pairs (i, j) as body indices, 2 ints each
    int* npairs;
    int cap;                 // pair capacity
    int reset;               // compute the interaction accelerations first
    int* fault;
};

// Changeover weight of mfo_drct (mercury6_2.for:4735-4760) times 1/s^3.
__device__ __forceinline__ double drct_factor(double s2, double rc) {
    double rc2 = rc * rc;
    if (s2 >= rc2) {
        double s_1 = 1.0 / sqrt(s2);
        return s_1 * s_1 * s_1;
    }
    if (s2 <= 0.01 * rc2) return 0.0;
    double s_1 = 1.0 / sqrt(s2);
    double s = 1.0 / s_1;
    double s_3 = s_1 * s_1 * s_1;
    double q = (s - 0.1 * rc) / (0.9 * rc);
    double q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3;
    return (10.0 * q3 - 15.0 * q4 + 6.0 * q5) * s_3;
}

// mfo_hy for the Big bodies at positions X (LDS), lane p = body p+1: the
// Fortran pair order (earlier partners subtract, later ones add; massless
// partners contribute zero), then oblateness; lane 0 writes acen.
__device__ inline void hy_big_accel(const Sys& d, int np, const double* m, const double* rc, const double* X,
                                    double* A, double* aob, double* acen) {
    const int p = threadIdx.x;
    if (p < np) {
        const double* xb = X + 3 * p;
        double a[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < p; i++) {
            double dx = xb[0] - X[3 * i], dy = xb[1] - X[3 * i + 1], dz = xb[2] - X[3 * i + 2];
            double tmp2 = drct_factor(dx * dx + dy * dy + dz * dz, fmax(rc[i], rc[p]));
            double faci = tmp2 * m[i + 1];
            a[0] = a[0] - faci * dx; a[1] = a[1] - faci * dy; a[2] = a[2] - faci * dz;
        }
        for (int j = p + 1; j < np; j++) {
            double dx = X[3 * j] - xb[0], dy = X[3 * j + 1] - xb[1], dz = X[3 * j + 2] - xb[2];
            double tmp2 = drct_factor(dx * dx + dy * dy + dz * dz, fmax(rc[p], rc[j]));
            double facj = tmp2 * m[j + 1];
            a[0] = a[0] + facj * dx; a[1] = a[1] + facj * dy; a[2] = a[2] + facj * dz;
        }
        for (int c = 0; c < 3; c++) A[3 * p + c] = a[c];
        if (d.obl) obl_body(d.mu, d.j2, d.j4, d.j6, xb[0], xb[1], xb[2], aob[3 * p], aob[3 * p + 1], aob[3 * p + 2]);
    }
    __syncthreads();
    if (p == 0) {
        double ac[3] = {0.0, 0.0, 0.0};
        if (d.obl)
            for (int q = 0; q < np; q++) {
                double tmp4 = m[q + 1] / m[0];
                for (int c = 0; c < 3; c++) ac[c] = ac[c] - tmp4 * aob[3 * q + c];
            }
        for (int c = 0; c < 3; c++) acen[c] = ac[c];
    }
    __syncthreads();
    if (p < np && d.obl)
        for (int c = 0; c < 3; c++) A[3 * p + c] = A[3 * p + c] + aob[3 * p + c] - acen[c];
    __syncthreads();
}

// Momentum sum of mdt_hy / mco_dh2h over the Big bodies in body order (massless
// particles add exact zeros), times temp.
__device__ inline void momentum_shift(int np, const double* m, const double* V, double temp, double* out) {
    double mv[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < np; j++)
        for (int c = 0; c < 3; c++) mv[c] = mv[c] + m[j + 1] * V[3 * j + c];
    for (int c = 0; c < 3; c++) out[c] = temp * mv[c];
}

__device__ inline void push_pair(const HyArgs& a, int i, int j) {
    int k = atomicAdd(a.npairs, 1);
    if (k < a.cap) {
        a.pairs[2 * k] = i;
        a.pairs[2 * k + 1] = j;
    } else {
        atomicOr(a.fault, FAULT_REQ_OVERFLOW);
    }
}

// mce_snif's test for one pair (mercury6_2.for:1700-1748), i the Big body.
__device__ __forceinline__ bool snif_pair(double h, const double* xi0, const double* vi0, const double* xi1,
                                         const double* vi1, const double* xj0, const double* vj0, const double* xj1,
                                         const double* vj1, double rc) {
    double rc2 = rc * rc;
    double dx0 = xi0[0] - xj0[0], dy0 = xi0[1] - xj0[1], dz0 = xi0[2] - xj0[2];
    double dx1 = xi1[0] - xj1[0], dy1 = xi1[1] - xj1[1], dz1 = xi1[2] - xj1[2];
    double d0 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0;
    double d1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
    double du0 = vi0[0] - vj0[0], dv0 = vi0[1] - vj0[1], dw0 = vi0[2] - vj0[2];
    double du1 = vi1[0] - vj1[0], dv1 = vi1[1] - vj1[1], dw1 = vi1[2] - vj1[2];
    double d0t = (dx0 * du0 + dy0 * dv0 + dz0 * dw0) * 2.0;
    double d1t = (dx1 * du1 + dy1 * dv1 + dz1 * dw1) * 2.0;
    double d2min = 9.9e29, tmin;
    if (d0t * h <= 0 && d1t * h >= 0) mce_min(d0, d1, d0t, d1t, h, d2min, tmin);
    double temp = fmin(fmin(d0, d1), d2min);
    return temp <= rc2;
}

__global__ void __launch_bounds__(64) hy_big_begin(HyArgs args) {
    const Sys d = args.sys[0];
    const int np = d.nbig - 1;
    const int p = threadIdx.x;
    const long long cap = args.st.cap;
    const int b0 = d.base;
    const double h = d.h, hby2 = d.hby2;
    __shared__ double m[kMaxBig], rc[kMaxBig], X[3 * kMaxBig], V[3 * kMaxBig], A[3 * kMaxBig], aob[3 * kMaxBig];
    __shared__ double hdr[kHyHdr];
    __shared__ double X0[3 * kMaxBig], V0[3 * kMaxBig], X1[3 * kMaxBig], V1[3 * kMaxBig], BOX[4 * kMaxBig];
    const bool ng = d.ngflag == 1 || d.ngflag == 3;
    for (int q = p; q < d.nbig; q += 64) m[q] = args.st.m[b0 + q];
    for (int q = p; q < np; q += 64) {
        rc[q] = args.rcrit[b0 + 1 + q];
        for (int c = 0; c < 3; c++) {
            X[3 * q + c] = args.st.x[c * cap + b0 + 1 + q];
            V[3 * q + c] = args.st.v[c * cap + b0 + 1 + q];
            A[3 * q + c] = args.st.a[c * cap + b0 + 1 + q];
        }
    }
    __syncthreads();
    double* T = args.tab;
    double ang[3] = {0.0, 0.0, 0.0};
    if (p < np) {
        for (int c = 0; c < 3; c++) T[kHyHdr + HB_N * p + HB_XS + c] = X[3 * p + c];
        if (ng && args.reset) {
            double pb[5];
            load_ngf(args.st, b0 + 1 + p, pb);
            ngf_body(pb[0], pb[1], pb[2], pb[4], X[3 * p], X[3 * p + 1], X[3 * p + 2], V[3 * p], V[3 * p + 1],
                     V[3 * p + 2], ang[0], ang[1], ang[2], args.fault);
            for (int c = 0; c < 3; c++) args.st.angf[c * cap + b0 + 1 + p] = ang[c];
        } else if (ng) {
            for (int c = 0; c < 3; c++) ang[c] = args.st.angf[c * cap + b0 + 1 + p];
        }
    }
    if (args.reset) hy_big_accel(d, np, m, rc, X, A, aob, hdr + 6);
    else if (p == 0) for (int c = 0; c < 3; c++) hdr[6 + c] = 0.0;
    if (p == 0) {
        momentum_shift(np, m, V, 1.0 / m[0], hdr);     // physical velocity of the start state
    }
    __syncthreads();
    if (p < np)
        for (int c = 0; c < 3; c++) {
            args.st.ox[c * cap + b0 + 1 + p] = X[3 * p + c];
            args.st.ov[c * cap + b0 + 1 + p] = V[3 * p + c] + hdr[c];
            args.st.a[c * cap + b0 + 1 + p] = A[3 * p + c];
            V[3 * p + c] = V[3 * p + c] + hby2 * (ang[c] + 0.0 + A[3 * p + c]);
        }
    __syncthreads();
    if (p == 0) momentum_shift(np, m, V, hby2 / m[0], hdr + 3);
    __syncthreads();
    if (p < np) {
        for (int c = 0; c < 3; c++) {
            X[3 * p + c] = X[3 * p + c] + hdr[3 + c];
            X0[3 * p + c] = X[3 * p + c];
            V0[3 * p + c] = V[3 * p + c];
        }
        Orbit o{X[3 * p], X[3 * p + 1], X[3 * p + 2], V[3 * p], V[3 * p + 1], V[3 * p + 2]};
        drift_one(m[0], o, h);
        X1[3 * p] = o.x; X1[3 * p + 1] = o.y; X1[3 * p + 2] = o.z;
        V1[3 * p] = o.vx; V1[3 * p + 1] = o.vy; V1[3 * p + 2] = o.vz;
        double lo, hi;
        box_axis(h, X0[3 * p], V0[3 * p], X1[3 * p], V1[3 * p], lo, hi);
        BOX[4 * p] = lo - rc[p]; BOX[4 * p + 1] = hi + rc[p];
        box_axis(h, X0[3 * p + 1], V0[3 * p + 1], X1[3 * p + 1], V1[3 * p + 1], lo, hi);
        BOX[4 * p + 2] = lo - rc[p]; BOX[4 * p + 3] = hi + rc[p];
        double* B = T + kHyHdr + HB_N * p;
        for (int c = 0; c < 3; c++) {
            B[HB_X0 + c] = X0[3 * p + c]; B[HB_V0 + c] = V0[3 * p + c];
            B[HB_X1 + c] = X1[3 * p + c]; B[HB_V1 + c] = V1[3 * p + c];
            args.x0[c * cap + b0 + 1 + p] = X0[3 * p + c];
            args.v0[c * cap + b0 + 1 + p] = V0[3 * p + c];
            args.st.x[c * cap + b0 + 1 + p] = X1[3 * p + c];
            args.st.v[c * cap + b0 + 1 + p] = V1[3 * p + c];
        }
        for (int k = 0; k < 4; k++) B[HB_BOX + k] = BOX[4 * p + k];
        B[HB_RC] = rc[p];
        B[HB_M] = m[p + 1];
    }
    if (p < kHyHdr) T[p] = p < 9 ? hdr[p] : 0.0;
    __syncthreads();
    // Big-Big pairs (both boxes padded)
    if (p < np)
        for (int j = p + 1; j < np; j++) {
            if (!(BOX[4 * p + 1] >= BOX[4 * j] && BOX[4 * j + 1] >= BOX[4 * p] && BOX[4 * p + 3] >= BOX[4 * j + 2]
                  && BOX[4 * j + 3] >= BOX[4 * p + 2]))
                continue;
            if (snif_pair(h, X0 + 3 * p, V0 + 3 * p, X1 + 3 * p, V1 + 3 * p, X0 + 3 * j, V0 + 3 * j, X1 + 3 * j,
                          V1 + 3 * j, fmax(rc[p], rc[j])))
                push_pair(args, b0 + 1 + p, b0 + 1 + j);
        }
}

// mfo_hy acceleration of a massless particle from the Big bodies at Bx.
template <int STRIDE>
__device__ inline void hy_particle_accel(const Sys& d, const double* __restrict__ T, int np, int off, double rcj,
                                         const double* x, const double* acen, double* a) {
    a[0] = 0.0; a[1] = 0.0; a[2] = 0.0;
    for (int i = 0; i < np; i++) {
        const double* B = T + kHyHdr + STRIDE * i;
        double dx = x[0] - B[off], dy = x[1] - B[off + 1], dz = x[2] - B[off + 2];
        double tmp2 = drct_factor(dx * dx + dy * dy + dz * dz, fmax(B[HB_RC], rcj));
        double faci = tmp2 * B[HB_M];
        a[0] = a[0] - faci * dx; a[1] = a[1] - faci * dy; a[2] = a[2] - faci * dz;
    }
    if (d.obl) {
        double ob[3];
        obl_body(d.mu, d.j2, d.j4, d.j6, x[0], x[1], x[2], ob[0], ob[1], ob[2]);
        for (int c = 0; c < 3; c++) a[c] = a[c] + ob[c] - acen[c];
    }
}

template <bool NG>
__global__ void __launch_bounds__(128) hy_particles_begin(HyArgs args) {
    const Sys d = args.sys[0];
    const int jb = d.nbig + blockIdx.x * blockDim.x + threadIdx.x;
    if (jb >= d.n) return;
    const int np = d.nbig - 1;
    const long long cap = args.st.cap;
    const int g = d.base + jb;
    const double* T = args.tab;
    const double h = d.h, hby2 = d.hby2;
    double x[3], v[3], a[3];
    for (int c = 0; c < 3; c++) {
        x[c] = args.st.x[c * cap + g];
        v[c] = args.st.v[c * cap + g];
    }
    const double rcj = args.rcrit[g];
    if (args.reset) hy_particle_accel<HB_N>(d, T, np, HB_XS, rcj, x, T + 6, a);
    else for (int c = 0; c < 3; c++) a[c] = args.st.a[c * cap + g];
    double ang[3] = {0.0, 0.0, 0.0};
    if (NG && args.reset) {
        double pb[5];
        load_ngf(args.st, g, pb);
        ngf_body(pb[0], pb[1], pb[2], pb[4], x[0], x[1], x[2], v[0], v[1], v[2], ang[0], ang[1], ang[2],
                 args.fault);
        for (int c = 0; c < 3; c++) args.st.angf[c * cap + g] = ang[c];
    } else if (NG) {
        for (int c = 0; c < 3; c++) ang[c] = args.st.angf[c * cap + g];
    }
    double x0[3], v0[3];
    for (int c = 0; c < 3; c++) {
        args.st.ox[c * cap + g] = x[c];
        args.st.ov[c * cap + g] = v[c] + T[c];
        if (args.reset) args.st.a[c * cap + g] = a[c];
        v[c] = v[c] + hby2 * (ang[c] + 0.0 + a[c]);
        x[c] = x[c] + T[3 + c];
        x0[c] = x[c];
        v0[c] = v[c];
        args.x0[c * cap + g] = x0[c];
        args.v0[c * cap + g] = v0[c];
    }
    Orbit o{x[0], x[1], x[2], v[0], v[1], v[2]};
    drift_one(d.mu, o, h);
    double x1[3] = {o.x, o.y, o.z}, v1[3] = {o.vx, o.vy, o.vz};
    for (int c = 0; c < 3; c++) {
        args.st.x[c * cap + g] = x1[c];
        args.st.v[c * cap + g] = v1[c];
    }
    double xmin, xmax, ymin, ymax;
    box_axis(h, x0[0], v0[0], x1[0], v1[0], xmin, xmax);
    box_axis(h, x0[1], v0[1], x1[1], v1[1], ymin, ymax);
    for (int i = 0; i < np; i++) {
        const double* B = T + kHyHdr + HB_N * i;
        if (!(B[HB_BOX + 1] >= xmin && xmax >= B[HB_BOX] && B[HB_BOX + 3] >= ymin && ymax >= B[HB_BOX + 2])) continue;
        if (snif_pair(h, B + HB_X0, B + HB_V0, B + HB_X1, B + HB_V1, x0, v0, x1, v1, fmax(B[HB_RC], rcj)))
            push_pair(args, d.base + 1 + i, g);
    }
}

// Encounter members back to their pre-drift states, gathered for the host.
__global__ void hy_revert_gather(State st, const double* x0, const double* v0, const int* ids, int count,
                                 double* packed) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= count) return;
    const long long cap = st.cap;
    const int j = ids[a];
    for (int c = 0; c < 3; c++) {
        st.x[c * cap + j] = x0[c * cap + j];
        st.v[c * cap + j] = v0[c * cap + j];
        packed[6 * a + c] = x0[c * cap + j];
        packed[6 * a + 3 + c] = v0[c * cap + j];
    }
}

__global__ void hy_scatter(State st, const int* ids, int count, const double* packed) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= count) return;
    const long long cap = st.cap;
    const int j = ids[a];
    for (int c = 0; c < 3; c++) {
        st.x[c * cap + j] = packed[6 * a + c];
        st.v[c * cap + j] = packed[6 * a + 3 + c];
    }
}

// A per-substep BS2 attempt's bookkeeping: the fault flags and the 12 level
// errors cleared before it, and reported after it into the host's pinned
// words (out[0..11] the errors' bits, out[12] the fault flags).
__global__ void bs_attempt_clear(int* fault, unsigned long long* err) {
    const int t = threadIdx.x;
    if (t < 12) err[t] = 0ull;
    if (t == 12) *fault = 0;
}

__global__ void bs_attempt_report(const unsigned long long* err, const int* fault, unsigned long long* out) {
    const int t = threadIdx.x;
    if (t < 12) out[t] = err[t];
    if (t == 12) out[12] = (unsigned long long)(unsigned)*fault;
}

// Encounter members' x, v ([3][ncap] each) in the driver's layout, x(3,n)
// then v(3,n); out is the device address of the host's pinned staging.
__global__ void enc_pack_aos(const double* x, const double* v, int n, int ncap, double* out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int c = 0; c < 3; c++) {
        out[3 * j + c] = x[(long long)c * ncap + j];
        out[(long long)3 * n + 3 * j + c] = v[(long long)c * ncap + j];
    }
}

// Finish table: header mvs (solar shift), mve (physical shift of the end
// state), acen; per Big: x after the shift (HF_X), rcrit, mass.
enum { HF_X = 0, HF_N = 21 };

__global__ void __launch_bounds__(64) hy_big_finish(HyArgs args) {
    const Sys d = args.sys[0];
    const int np = d.nbig - 1;
    const int p = threadIdx.x;
    const long long cap = args.st.cap;
    const int b0 = d.base;
    const double hby2 = d.hby2;
    __shared__ double m[kMaxBig], rc[kMaxBig], X[3 * kMaxBig], V[3 * kMaxBig], A[3 * kMaxBig], aob[3 * kMaxBig];
    __shared__ double hdr[kHyHdr];
    for (int q = p; q < d.nbig; q += 64) m[q] = args.st.m[b0 + q];
    for (int q = p; q < np; q += 64) {
        rc[q] = args.rcrit[b0 + 1 + q];
        for (int c = 0; c < 3; c++) {
            X[3 * q + c] = args.st.x[c * cap + b0 + 1 + q];
            V[3 * q + c] = args.st.v[c * cap + b0 + 1 + q];
        }
    }
    __syncthreads();
    if (p == 0) momentum_shift(np, m, V, hby2 / m[0], hdr);
    __syncthreads();
    if (p < np)
        for (int c = 0; c < 3; c++) X[3 * p + c] = X[3 * p + c] + hdr[c];
    __syncthreads();
    hy_big_accel(d, np, m, rc, X, A, aob, hdr + 6);
    if (p < np) {
        double ang[3] = {0.0, 0.0, 0.0};
        if (d.ngflag == 1 || d.ngflag == 3) {
            double pb[5];
            load_ngf(args.st, b0 + 1 + p, pb);
            ngf_body(pb[0], pb[1], pb[2], pb[4], X[3 * p], X[3 * p + 1], X[3 * p + 2], V[3 * p], V[3 * p + 1],
                     V[3 * p + 2], ang[0], ang[1], ang[2], args.fault);
            for (int c = 0; c < 3; c++) args.st.angf[c * cap + b0 + 1 + p] = ang[c];
        }
        for (int c = 0; c < 3; c++) V[3 * p + c] = V[3 * p + c] + hby2 * (ang[c] + 0.0 + A[3 * p + c]);
    }
    __syncthreads();
    if (p == 0) momentum_shift(np, m, V, 1.0 / m[0], hdr + 3);
    __syncthreads();
    double* T = args.tab;
    if (p < np) {
        double* B = T + kHyHdr + HF_N * p;
        for (int c = 0; c < 3; c++) {
            B[HF_X + c] = X[3 * p + c];
            args.st.x[c * cap + b0 + 1 + p] = X[3 * p + c];
            args.st.v[c * cap + b0 + 1 + p] = V[3 * p + c];
            args.st.a[c * cap + b0 + 1 + p] = A[3 * p + c];
            args.st.v1[c * cap + b0 + 1 + p] = V[3 * p + c] + hdr[3 + c];
        }
        B[HB_RC] = rc[p];
        B[HB_M] = m[p + 1];
    }
    if (p < kHyHdr) T[p] = p < 9 ? hdr[p] : 0.0;
}

template <bool NG>
__global__ void __launch_bounds__(128) hy_particles_finish(HyArgs args) {
    const Sys d = args.sys[0];
    const int jb = d.nbig + blockIdx.x * blockDim.x + threadIdx.x;
    if (jb >= d.n) return;
    const int np = d.nbig - 1;
    const long long cap = args.st.cap;
    const int g = d.base + jb;
    const double* T = args.tab;
    double x[3], a[3];
    for (int c = 0; c < 3; c++) x[c] = args.st.x[c * cap + g] + T[c];
    hy_particle_accel<HF_N>(d, T, np, HF_X, args.rcrit[g], x, T + 6, a);
    double ang[3] = {0.0, 0.0, 0.0};
    if (NG) {
        double pb[5];
        load_ngf(args.st, g, pb);
        ngf_body(pb[0], pb[1], pb[2], pb[4], x[0], x[1], x[2], args.st.v[g], args.st.v[cap + g],
                 args.st.v[2 * cap + g], ang[0], ang[1], ang[2], args.fault);
        for (int c = 0; c < 3; c++) args.st.angf[c * cap + g] = ang[c];
    }
    bool bad = false;
    for (int c = 0; c < 3; c++) {
        double v = args.st.v[c * cap + g] + d.hby2 * (ang[c] + 0.0 + a[c]);
        args.st.x[c * cap + g] = x[c];
        args.st.v[c * cap + g] = v;
        args.st.a[c * cap + g] = a[c];
        args.st.v1[c * cap + g] = v + T[3 + c];
        bad = bad || !isfinite(a[c]);
    }
    if (bad) atomicOr(args.fault, FAULT_FORCE);
}

// ---------------------------------------------------------------- compact BS2
constexpr int kEncBig = 16;     // Big members of one encounter set
constexpr int kBsPoints = 78;   // force points of the 12 modified-midpoint levels

// Step control of the driver's mdt_hkce loop (mercury6_2.for:3883-3905) with
// mdt_bs2's step rule (mercury6_2.for:3428-3617) on the device, for callers
// that do not list encounters after every substep (bs2_loop).
struct BsCtl {
    double h0;                  // the step the encounter set is integrated over
    double tlocal;              // time done so far
    double hlocal;              // step of the next attempt
    double hrec;                // mdt_hkce's recommendation for the next encounter
    double tol2;
    long long forces, rejected, substeps;
    int state;                  // 0 running, 1 done, 2 step below the time's resolution, 3 non-finite
};

struct BsArgs {
    int n, nbig;                // members (index 0: the central body), Big members 1..nbig-1
    const double* m;            // [n]
    const double* rcrit;        // [n]
    double* x;                  // [3][ncap] state at the attempt's start
    double* v;
    int ncap;
    const int* bbp;             // Big-Big pairs (compact indices, 1..nbig-1), list order
    int nbbp;
    const int* poff;            // particle partners: CSR offsets [n - nbig + 1]
    const int* pidx;            // partner Big members (compact index)
    const int4* pinfo;          // per particle: partners held in registers (or -1), their indices
    const double4* pcon;        // per particle: their max(rcrit) and masses (pack_partners)
    double h0, tol2;
    double* tp;                 // Big positions at every force point [kBsPoints][3 * kEncBig] (bs2_big)
    double* cand;               // accepted states per level n = 4..12: [9][6][ncap]
    unsigned long long* err;    // max error per level [12] (bits of non-negative doubles)
    int* fault;
};

// mfo_hkce pair factor (1 - K) / s^3 inside rc, else 0 (mercury6_2.for, mfo_hkce).
template <class Ops>
__device__ __forceinline__ void hkce_factor_ops(double s2, double rc, double rc2, double& tmp2) {
    double s_1 = Ops::rcp(Ops::sqrt(s2));
    double s_3 = s_1 * s_1 * s_1;
    if (s2 <= 0.01 * rc2) {
        tmp2 = s_3;
    } else {
        double s = Ops::rcp(s_1);
        double q = Ops::div(s - 0.1 * rc, 0.9 * rc);
        double q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3;
        tmp2 = (1.0 - 10.0 * q3 + 15.0 * q4 - 6.0 * q5) * s_3;
    }
}

// With rc in [2^-200, 2^200] and s^2 inside RangeOps' root range (sq_ok), the
// short exact sequences of rn_ops.hpp give the IEEE bits: s^-1 and s lie in
// [2^-500, 2^384]; in the outer branch s - 0.1 rc is zero or at least a few
// ulp of 0.1 rc (>= 2^-256) and at most rc, and 0.9 rc is in range, so the
// quotient's operands and result lie where div_rn is exact.  Outside the gate
// the IEEE forms run.
__device__ __forceinline__ bool hkce_factor(double s2, double rc, double& tmp2) {
    double rc2 = rc * rc;
    if (!(s2 < rc2)) return false;
    if (RangeOps::sq_ok(s2) && rc >= 0x1p-200 && rc <= 0x1p200) hkce_factor_ops<RangeOps>(s2, rc, rc2, tmp2);
    else hkce_factor_ops<IeeeOps>(s2, rc, rc2, tmp2);
    return true;
}

// 1 / sqrt(s2) with the IEEE bits: the short exact sequences where RangeOps
// admits s2, the IEEE forms elsewhere.
__device__ __forceinline__ double rsqrt_exact(double s2) {
    if (RangeOps::sq_ok(s2)) return RangeOps::rcp(RangeOps::sqrt(s2));
    return 1.0 / sqrt(s2);
}

__device__ __forceinline__ void merge_err(unsigned long long* slot, double e) {
    unsigned long long bits = (unsigned long long)__double_as_longlong(e);
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(bits, off);
        bits = o > bits ? o : bits;
    }
    if (__lane_id() == 0) atomicMax(slot, bits);
}

// The Big members' pair list, rcrit and masses, staged in LDS once per
// launch: the force rounds then read no global memory between barriers.
constexpr int kBbMax = 256;     // Big-Big pairs one encounter set may list
struct BigShared {
    int bb[2 * kBbMax];
    double rc[kEncBig + 1], m[kEncBig + 1];
};

// Threads t of nt copy the lists; a barrier must follow before big_accel.
__device__ __forceinline__ void stage_big(const BsArgs& a, BigShared& S, int t, int nt) {
    for (int w = t; w < 2 * a.nbbp; w += nt) S.bb[w] = a.bbp[w];
    for (int w = t; w < a.nbig; w += nt) { S.rc[w] = a.rcrit[w]; S.m[w] = a.m[w]; }
}

// Force on Big member me (1..nb) from the Big positions P ([3][kEncBig]):
// mfo_hkce's Big-Big pairs in list order, then the central term.
__device__ inline void big_accel(const BigShared& S, int nbbp, int me, const double* P, double* acc) {
    acc[0] = 0.0; acc[1] = 0.0; acc[2] = 0.0;
    for (int q = 0; q < nbbp; q++) {
        int i = S.bb[2 * q], j = S.bb[2 * q + 1];
        if (i != me && j != me) continue;
        double dx = P[3 * (j - 1)] - P[3 * (i - 1)];
        double dy = P[3 * (j - 1) + 1] - P[3 * (i - 1) + 1];
        double dz = P[3 * (j - 1) + 2] - P[3 * (i - 1) + 2];
        double tmp2;
        if (!hkce_factor(dx * dx + dy * dy + dz * dz, fmax(S.rc[i], S.rc[j]), tmp2)) continue;
        if (j == me) {
            double faci = tmp2 * S.m[i];
            acc[0] = acc[0] - faci * dx; acc[1] = acc[1] - faci * dy; acc[2] = acc[2] - faci * dz;
        } else {
            double facj = tmp2 * S.m[j];
            acc[0] = acc[0] + facj * dx; acc[1] = acc[1] + facj * dy; acc[2] = acc[2] + facj * dz;
        }
    }
    const double* x = P + 3 * (me - 1);
    double s2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    double s_1 = rsqrt_exact(s2);
    double tmp2 = S.m[0] * s_1 * s_1 * s_1;
    for (int c = 0; c < 3; c++) acc[c] = acc[c] - tmp2 * x[c];
}

__device__ __forceinline__ void bs_scales(const double* x0, const double* v0, double& xs, double& vs) {
    double t1 = x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2];
    double t2 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
    xs = 1.0 / fmax(t1, 1e-300);
    vs = 1.0 / fmax(t2, 1e-300);
}

// mdt_bs2's 12 modified-midpoint levels (mercury6_2.for:3470-3530) are
// independent until the extrapolation, so each level runs in its own thread;
// level n makes its n force evaluations at the force points n(n-1)/2 ..
// n(n+1)/2 - 1 of the sequential order.  bs_level_* follow one level of one
// member: start, then per evaluation the update, then the level's result.
constexpr int kLevels = 12;
__host__ __device__ constexpr int bs_point(int n, int r) { return n * (n - 1) / 2 + r - 1; }   // r = 1..n

struct BsLevel {
    double h, hby2, h2, h2by2, tmp0;
    double b[3], c[3], xend[3];
};

__device__ __forceinline__ void bs_level_start(BsLevel& L, int n, double h0, const double* x0, const double* v0,
                                               const double* a0) {
    L.h = h0 / (double)n;
    L.hby2 = .5 * L.h;
    L.h2 = L.h * L.h;
    L.h2by2 = .5 * L.h2;
    for (int q = 0; q < 3; q++) {
        L.b[q] = .5 * a0[q];
        L.c[q] = 0.0;
        L.xend[q] = L.h2by2 * a0[q] + L.h * v0[q] + x0[q];
    }
}

// After the force a at evaluation r of n (r < n): advance xend to the next point.
__device__ __forceinline__ void bs_level_step(BsLevel& L, int r, const double* a, const double* x0,
                                              const double* v0, const double* a0) {
    const int j = r + 1;
    double tmp0 = L.h * (double)j;
    for (int q = 0; q < 3; q++) {
        L.b[q] = L.b[q] + a[q];
        L.c[q] = L.c[q] + L.b[q];
        L.xend[q] = L.h2 * L.c[q] + L.h2by2 * a0[q] + tmp0 * v0[q] + x0[q];
    }
}

// The level's result from the force a at its last evaluation: D row n-1.
__device__ __forceinline__ void bs_level_result(const BsLevel& L, const double* a, const double* v0, double* R) {
    for (int q = 0; q < 3; q++) {
        R[q] = L.xend[q];
        R[3 + q] = L.h * L.b[q] + L.hby2 * a[q] + v0[q];
    }
}

// mdt_bs2's extrapolation weights for the step h0 (mercury6_2.for:3533-3536):
// for level n and j = n-1 .. 1, tmp0 = 1/(h2(j) - h2(n)), W = (tmp0 * h2(j+1),
// tmp0 * h2(n)), with h2(n) = (h0/n)^2.  They depend on h0 alone, so a block
// forms them once (threads t of nt; a barrier must follow) for all its members.
constexpr int kExPairs = kLevels * (kLevels - 1) / 2;
__host__ __device__ constexpr int ex_index(int n, int j) { return (n - 1) * (n - 2) / 2 + j - 1; }
__device__ inline void ex_weights(double h0, double (*W)[2], int t, int nt) {
    for (int w = t; w < kExPairs; w += nt) {
        int n = 2;
        while (n * (n - 1) / 2 <= w) n++;
        const int j = w - (n - 1) * (n - 2) / 2 + 1;
        const double hn = h0 / (double)n, hj = h0 / (double)j, hj1 = h0 / (double)(j + 1);
        const double h2n = hn * hn, h2j = hj * hj, h2j1 = hj1 * hj1;
        const double tmp0 = 1.0 / (h2j - h2n);
        W[w][0] = tmp0 * h2j1;
        W[w][1] = tmp0 * h2n;
    }
}

// Extrapolation of one component q of one member over the levels in
// mdt_bs2's order (mercury6_2.for:3531-3570), given every level's result
// already in D (row n-1 holds level n's: no earlier level writes it) and the
// weights W.  Per level n > 3 it keeps d(1) in E[n-4] and writes the accepted
// value.
__device__ inline void bs_extrapolate(const double (*W)[2], double* D, int ds, double* E, double* cand,
                                      long long cs) {
    for (int n = 1; n <= kLevels; n++) {
        for (int j = n - 1; j >= 1; j--) {
            const int j1 = j + 1;
            const double tmp1 = W[ex_index(n, j)][0];
            const double tmp2 = W[ex_index(n, j)][1];
            D[(j - 1) * 6 * ds] = tmp1 * D[(j1 - 1) * 6 * ds] - tmp2 * D[(j - 1) * 6 * ds];
        }
        if (n > 3) {
            double d1 = D[0];
            E[n - 4] = d1;
            double s = d1;
            for (int j = 2; j <= n; j++) s = s + D[(j - 1) * 6 * ds];
            cand[(long long)(n - 4) * 6 * cs] = s;
        }
    }
}

// Level error of one member from its six d(1) values (mercury6_2.for:3544-3556).
__device__ __forceinline__ double bs_level_err(const double* d1, double xs, double vs, int* fault) {
    bool fin = true;
    for (int q = 0; q < 6; q++) fin = fin && isfinite(d1[q]);
    double tmp1 = fmax(fmax(d1[0] * d1[0], d1[1] * d1[1]), d1[2] * d1[2]);
    double tmp2 = fmax(fmax(d1[3] * d1[3], d1[4] * d1[4]), d1[5] * d1[5]);
    if (!fin) atomicOr(fault, FAULT_FORCE);
    return fin ? fmax(tmp1 * xs, tmp2 * vs) : 0.0;
}

// One BS2 attempt, Big members: one block of kLevels x kEncBig threads,
// thread (level n, member k+1).  All levels advance together, one force
// evaluation per round, with a block barrier between writing the positions
// and reading them for the pair forces; every position is also written to
// tp (the Big positions at each force point) for bs2_particles.  Then the
// per-component extrapolation, the level errors and the accepted states.
// The attempt starts from member j's state at xs[c * ncap + j], vs[c * ncap + j].
__device__ inline void bs2_big_body(const BsArgs& A, const double* xs, const double* vs) {
    const double h0 = A.h0;
    const int nb = A.nbig - 1;
    const int t = threadIdx.x;
    const int n = t / kEncBig + 1;
    const int k = t % kEncBig;
    const bool mem = k < nb;
    __shared__ double P[kLevels][3 * kEncBig];
    __shared__ double X0[3 * kEncBig];
    __shared__ double D[72 * kEncBig];       // [6 * level + component][member]
    __shared__ double E[kEncBig][6][9];
    __shared__ BigShared S;
    __shared__ double W[kExPairs][2];
    __shared__ double XS[kEncBig][2];        // the member's error scales
    stage_big(A, S, t, kLevels * kEncBig);
    ex_weights(h0, W, t, kLevels * kEncBig);
    double x0[3] = {0, 0, 0}, v0[3] = {0, 0, 0}, a0[3] = {0, 0, 0};
    if (mem) {
        for (int c = 0; c < 3; c++) {
            x0[c] = xs[c * A.ncap + 1 + k];
            v0[c] = vs[c * A.ncap + 1 + k];
            if (n == 1) X0[3 * k + c] = x0[c];
        }
        if (n == 1) bs_scales(x0, v0, XS[k][0], XS[k][1]);
    }
    __syncthreads();
    if (mem) big_accel(S, A.nbbp, k + 1, X0, a0);
    BsLevel L;
    bs_level_start(L, n, h0, x0, v0, a0);
    double a[3] = {0, 0, 0};
    for (int r = 1; r <= kLevels; r++) {
        const bool act = mem && r <= n;
        if (act)
            for (int c = 0; c < 3; c++) {
                P[n - 1][3 * k + c] = L.xend[c];
                A.tp[bs_point(n, r) * 3 * kEncBig + 3 * k + c] = L.xend[c];
            }
        __syncthreads();
        if (act) big_accel(S, A.nbbp, k + 1, P[n - 1], a);
        __syncthreads();
        if (act && r < n) bs_level_step(L, r, a, x0, v0, a0);
    }
    if (mem) {
        double R[6];
        bs_level_result(L, a, v0, R);
        for (int q = 0; q < 6; q++) D[(6 * (n - 1) + q) * kEncBig + k] = R[q];
    }
    __syncthreads();
    // extrapolation: thread (member k, component q)
    if (t < 6 * kEncBig) {
        const int km = t / 6, q = t % 6;
        if (km < nb) {
            double e[9];
            bs_extrapolate(W, D + q * kEncBig + km, kEncBig, e, A.cand + (long long)q * A.ncap + 1 + km, A.ncap);
            for (int l = 0; l < 9; l++) E[km][q][l] = e[l];
        }
    }
    __syncthreads();
    // level errors: thread (member k, level 4 + l)
    if (t < 9 * kEncBig) {
        const int km = t / 9, l = t % 9;
        if (km < nb) {
            double d1[6];
            for (int q = 0; q < 6; q++) d1[q] = E[km][q][l];
            double err = bs_level_err(d1, XS[km][0], XS[km][1], A.fault);
            atomicMax(A.err + 3 + l, (unsigned long long)__double_as_longlong(err));
        }
    }
}

__global__ void __launch_bounds__(kLevels * kEncBig) bs2_big(BsArgs A) { bs2_big_body(A, A.x, A.v); }

// One BS2 attempt, particle members (after bs2_big): kBsPairSlots particles
// per block.  Level n makes n force evaluations, so thread (pair p, slot) runs
// the levels p+1 and 12-p one after the other (13 evaluations each): no lane
// idles while a longer level of its wave runs, and a block holds twice the
// particles.  Each level's evaluations use the Big positions bs2_big recorded
// at the same force points; its arithmetic and d row are those of a thread of
// its own.  The d tables live in LDS; then the per-component extrapolation
// and the level errors, merged over the block before one atomicMax per level.
// The attempt starts from member j's state at xs[c * ncap + j], vs[c * ncap + j]
// (the state arrays, or the previous attempt's accepted candidates); RP holds
// the thread's particle's partners, loaded by the caller (load_partners).
constexpr int kBsPart = 16;                 // particle slots of bs2_loop_fused's blocks
constexpr int kBsPairSlots = 32;            // particles per block of bs2_particles
constexpr int kBsPairThreads = kLevels / 2 * kBsPairSlots;

// A particle's Big partners in registers -- the usual encounter set gives a
// particle one or two -- so that a force evaluation reads no pair list from
// memory: compact index, max(rcrit_i, rcrit_j) and mass, in list order.
// n = -1 when the particle has more partners (they are read from the lists).
struct RegPartners {
    int n;
    int i0, i1;
    double rc0, rc1, m0, m1;
};

__device__ __forceinline__ RegPartners reg_partners(const BsArgs& A, int p0, int p1, double rcj) {
    RegPartners R{p1 - p0 <= 2 ? p1 - p0 : -1, 1, 1, 0.0, 0.0, 0.0, 0.0};
    if (R.n > 0) { R.i0 = A.pidx[p0]; R.rc0 = fmax(A.rcrit[R.i0], rcj); R.m0 = A.m[R.i0]; }
    if (R.n > 1) { R.i1 = A.pidx[p0 + 1]; R.rc1 = fmax(A.rcrit[R.i1], rcj); R.m1 = A.m[R.i1]; }
    return R;
}

// Particle p's partners from the packed arrays pack_partners wrote.
__device__ __forceinline__ RegPartners load_partners(const BsArgs& A, int p) {
    const int4 f = A.pinfo[p];
    const double4 g = A.pcon[p];
    return RegPartners{f.x, f.y, f.z, g.x, g.y, g.z, g.w};
}

// Packs every particle's RegPartners once per encounter set: the pair lists
// and masses stay fixed while the set is integrated.
__global__ void pack_partners(BsArgs A, int4* info, double4* con) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n - A.nbig) return;
    const RegPartners R = reg_partners(A, A.poff[p], A.poff[p + 1], A.rcrit[A.nbig + p]);
    info[p] = make_int4(R.n, R.i0, R.i1, 0);
    con[p] = make_double4(R.rc0, R.rc1, R.m0, R.m1);
}

// One pair term of mfo_hkce for a particle at xe and a partner at X.
__device__ __forceinline__ void pair_term(const double* xe, const double* X, double rc, double m, double* acc) {
    double dx = xe[0] - X[0], dy = xe[1] - X[1], dz = xe[2] - X[2];
    double tmp2;
    if (!hkce_factor(dx * dx + dy * dy + dz * dz, rc, tmp2)) return;
    double faci = tmp2 * m;
    acc[0] = acc[0] - faci * dx; acc[1] = acc[1] - faci * dy; acc[2] = acc[2] - faci * dz;
}

// mfo_hkce for a particle whose partners are in R, at positions X0 and X1:
// the pair terms in list order, then the central term (mc = central mass),
// the same operations as the list loop.
__device__ __forceinline__ void pforce_reg(const RegPartners& R, const double* xe, const double* X0,
                                           const double* X1, double mc, double* out) {
    double acc[3] = {0.0, 0.0, 0.0};
    if (R.n > 0) pair_term(xe, X0, R.rc0, R.m0, acc);
    if (R.n > 1) pair_term(xe, X1, R.rc1, R.m1, acc);
    double s2 = xe[0] * xe[0] + xe[1] * xe[1] + xe[2] * xe[2];
    double s_1 = rsqrt_exact(s2);
    double tmp2 = mc * s_1 * s_1 * s_1;
    for (int c = 0; c < 3; c++) out[c] = acc[c] - tmp2 * xe[c];
}

// Partner positions at force point `pt` from bs2_big's table.
__device__ __forceinline__ void partner_positions(const RegPartners& R, const double* tp, int pt, double* X0,
                                                  double* X1) {
    const double* T = tp + pt * 3 * kEncBig;
    if (R.n > 0)
        for (int c = 0; c < 3; c++) X0[c] = T[3 * (R.i0 - 1) + c];
    if (R.n > 1)
        for (int c = 0; c < 3; c++) X1[c] = T[3 * (R.i1 - 1) + c];
}

__device__ inline void bs2_particles_body(const BsArgs& A, const double* xs, const double* vs,
                                         const RegPartners& RP) {
    constexpr int NS = kBsPairSlots;
    static_assert(6 * NS == kBsPairThreads, "one extrapolation item per thread");
    const double h0 = A.h0;
    const int t = threadIdx.x;
    const int pair = t / NS;
    const int sl = t % NS;
    const int np = A.n - A.nbig;
    const int pi_ = blockIdx.x * NS + sl;
    const bool live = pi_ < np;
    const int me = A.nbig + pi_;
    __shared__ double D[72 * NS];            // [6 * level + component][slot]
    __shared__ double E[NS][6][9];
    __shared__ double XS[NS][2];             // the slot's error scales
    __shared__ double W[kExPairs][2];
    __shared__ unsigned long long EM[9];
    if (t < 9) EM[t] = 0ull;
    ex_weights(h0, W, t, kBsPairThreads);
    double x0[3] = {1, 0, 0}, v0[3] = {0, 1, 0}, a0[3] = {0, 0, 0};
    double rcj = 0.0;
    int p0 = 0, p1 = 0;
    if (live) {
        for (int c = 0; c < 3; c++) { x0[c] = xs[c * A.ncap + me]; v0[c] = vs[c * A.ncap + me]; }
        if (RP.n < 0) {
            rcj = A.rcrit[me];
            p0 = A.poff[pi_];
            p1 = A.poff[pi_ + 1];
        }
        if (pair == 0) bs_scales(x0, v0, XS[sl][0], XS[sl][1]);
    }
    const double m0 = A.m[0];
    // mfo_hkce for a particle: pairs with its Big partners, then the central term;
    // Xb(i) gives Big member i's position at this evaluation
    auto pforce = [&](const double* xe, double* a, auto Xb) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int q = p0; q < p1; q++) {
            int i = A.pidx[q];
            const double* Xi = Xb(i);
            double dx = xe[0] - Xi[0], dy = xe[1] - Xi[1], dz = xe[2] - Xi[2];
            double tmp2;
            if (!hkce_factor(dx * dx + dy * dy + dz * dz, fmax(A.rcrit[i], rcj), tmp2)) continue;
            double faci = tmp2 * A.m[i];
            acc[0] = acc[0] - faci * dx; acc[1] = acc[1] - faci * dy; acc[2] = acc[2] - faci * dz;
        }
        double s2 = xe[0] * xe[0] + xe[1] * xe[1] + xe[2] * xe[2];
        double s_1 = rsqrt_exact(s2);
        double tmp2 = A.m[0] * s_1 * s_1 * s_1;
        for (int c = 0; c < 3; c++) a[c] = acc[c] - tmp2 * xe[c];
    };
    if (live) {
        if (RP.n >= 0) {
            double X0[3] = {0, 0, 0}, X1[3] = {0, 0, 0};
            for (int c = 0; c < 3; c++) {
                if (RP.n > 0) X0[c] = A.x[c * A.ncap + RP.i0];
                if (RP.n > 1) X1[c] = A.x[c * A.ncap + RP.i1];
            }
            pforce_reg(RP, x0, X0, X1, m0, a0);
        } else {
            double xb[3];
            pforce(x0, a0, [&](int i) {
                for (int c = 0; c < 3; c++) xb[c] = A.x[c * A.ncap + i];
                return (const double*)xb;
            });
        }
#pragma unroll 1
        for (int k = 0; k < 2; k++) {
            const int n = k == 0 ? pair + 1 : kLevels - pair;
            BsLevel L;
            bs_level_start(L, n, h0, x0, v0, a0);
            double a[3];
            if (RP.n >= 0) {
                // the next point's partner positions load while this one computes
                double X0[3] = {0, 0, 0}, X1[3] = {0, 0, 0}, Y0[3] = {0, 0, 0}, Y1[3] = {0, 0, 0};
                partner_positions(RP, A.tp, bs_point(n, 1), X0, X1);
                for (int r = 1; r <= n; r++) {
                    if (r < n) partner_positions(RP, A.tp, bs_point(n, r + 1), Y0, Y1);
                    pforce_reg(RP, L.xend, X0, X1, m0, a);
                    if (r < n) {
                        bs_level_step(L, r, a, x0, v0, a0);
                        for (int c = 0; c < 3; c++) { X0[c] = Y0[c]; X1[c] = Y1[c]; }
                    }
                }
            } else {
                for (int r = 1; r <= n; r++) {
                    const double* T = A.tp + bs_point(n, r) * 3 * kEncBig;
                    pforce(L.xend, a, [&](int i) { return T + 3 * (i - 1); });
                    if (r < n) bs_level_step(L, r, a, x0, v0, a0);
                }
            }
            double R[6];
            bs_level_result(L, a, v0, R);
            for (int q = 0; q < 6; q++) D[(6 * (n - 1) + q) * NS + sl] = R[q];
        }
    }
    __syncthreads();
    // extrapolation: thread (slot, component)
    {
        const int s = t / 6, q = t % 6;
        const int ps = blockIdx.x * NS + s;
        if (ps < np) {
            double e[9];
            bs_extrapolate(W, D + q * NS + s, NS, e, A.cand + (long long)q * A.ncap + A.nbig + ps, A.ncap);
            for (int l = 0; l < 9; l++) E[s][q][l] = e[l];
        }
    }
    __syncthreads();
    // level errors: item (slot, level 4 + l), merged per level over the block
    for (int w = t; w < 9 * NS; w += kBsPairThreads) {
        const int s = w / 9, l = w % 9;
        const int ps = blockIdx.x * NS + s;
        if (ps < np) {
            double d1[6];
            for (int q = 0; q < 6; q++) d1[q] = E[s][q][l];
            double err = bs_level_err(d1, XS[s][0], XS[s][1], A.fault);
            atomicMax(&EM[l], (unsigned long long)__double_as_longlong(err));
        }
    }
    __syncthreads();
    if (t < 9 && EM[t]) atomicMax(A.err + 3 + t, EM[t]);
}

__global__ void __launch_bounds__(kBsPairThreads) bs2_particles(BsArgs A) {
    const int p = blockIdx.x * kBsPairSlots + threadIdx.x % kBsPairSlots;
    const RegPartners RP = p < A.n - A.nbig ? load_partners(A, p) : RegPartners{0, 1, 1, 0.0, 0.0, 0.0, 0.0};
    bs2_particles_body(A, A.x, A.v, RP);
}

__global__ void bs2_accept(BsArgs A, int level) {
    if (level < 4) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k >= A.n) return;
    for (int q = 0; q < 3; q++) {
        A.x[q * A.ncap + k] = A.cand[(long long)((level - 4) * 6 + q) * A.ncap + k];
        A.v[q * A.ncap + k] = A.cand[(long long)((level - 4) * 6 + 3 + q) * A.ncap + k];
    }
}

// mdt_hkce's step control after one attempt (mercury6_2.for:3883-3905 with
// mdt_bs2's level choice, 3571-3605): take the first level n > 3 whose
// largest error is within tol^2, or halve the step; after an accepted substep
// advance the local time and, unless h0 is done, clamp the next step to what
// is left.  Returns the accepted level (0: none).  The arithmetic is the host
// loop's (bs2_step), operation for operation.
__device__ inline int bs_control_step(BsCtl& c, const unsigned long long* err, bool faulted) {
    const double SHRINK = .55, GROW = 1.3;
    if (faulted) { c.state = 3; return 0; }
    int level = 0;
    for (int n = 4; n <= 12 && !level; n++)
        if (__longlong_as_double((long long)err[n - 1]) <= c.tol2) level = n;
    if (level) {
        for (int n = 1; n <= level; n++) c.forces += n;
        const double hdid = c.hlocal;
        if (level >= 8) c.hlocal = c.hlocal * SHRINK;
        if (level < 7) c.hlocal = c.hlocal * GROW;
        c.tlocal = c.tlocal + hdid;
        c.substeps++;
        if (!((c.tlocal - c.h0) * c.h0 < 0.0)) { c.state = 1; return level; }
        c.forces += 1;                       // the next substep's start accelerations
        const double tmp0 = fabs(c.h0) - fabs(c.tlocal);
        c.hrec = c.hlocal;
        if (fabs(c.hlocal) > tmp0) c.hlocal = copysign(tmp0, c.h0);
    } else {
        c.forces += 78;
        c.rejected++;
        c.hlocal = c.hlocal * .5;
    }
    if (c.tlocal + c.hlocal == c.tlocal) c.state = 2;
    return level;
}

// Rotating buffers of the device encounter loop.  The state is stored, and
// the next error set cleared, by bs2_loop_big or block 0 of bs2_loop_fused.
struct BsLoop {
    BsCtl* ctl;                  // [2]: attempt q reads ctl[q & 1] and stores ctl[(q + 1) & 1]
    unsigned long long* err;     // [3][12]: attempt q accumulates into err[q % 3], clears err[(q + 1) % 3]
    double* cand;                // [2][9][6][ncap]: attempt q writes cand[q & 1]
};

// The whole mdt_hkce loop runs as two launches per BS2 attempt q, queued in
// batches with no host synchronisation: bs2_loop_big (one block), then
// bs2_loop_particles.  Each first makes attempt q-1's control decision from
// the same inputs -- the state attempt q-1 ran with (ctl[q & 1]), its level
// errors (err[(q - 1) % 3]) and the fault flag -- so every block of both
// kernels takes the same one, and applies its accept to its own members (the
// Big members, or the block's particles); bs2_loop_big alone stores the new
// state.  Then each runs attempt q exactly as bs2_big / bs2_particles do.
// Once a decision ends the loop, later launches only carry the state forward.
// Every input of the decision is loaded up front, in one round trip: the
// state, the 12 level errors of attempt q-1 (err[(q + 2) % 3], a valid set
// also for q = 0) and the fault flag; a loop that stopped at the first
// accepted level would wait on one load per level.
__device__ inline int bs_loop_decide(const BsArgs& A, const BsLoop& B, int q, BsCtl& c) {
    c = B.ctl[q & 1];
    const ulonglong2* ep = reinterpret_cast<const ulonglong2*>(B.err + 12 * ((q + 2) % 3));
    unsigned long long ev[12];
    for (int w = 0; w < 6; w++) {
        const ulonglong2 e2 = ep[w];
        ev[2 * w] = e2.x;
        ev[2 * w + 1] = e2.y;
    }
    const int fault = *A.fault;
    // an empty asm that takes every loaded value keeps the compiler from
    // sinking the loads behind the branches below
    unsigned long long all = (unsigned long long)fault ^ (unsigned long long)c.state ^
                             (unsigned long long)__double_as_longlong(c.hlocal) ^
                             (unsigned long long)__double_as_longlong(c.tlocal) ^
                             (unsigned long long)__double_as_longlong(c.tol2);
    for (int w = 0; w < 12; w++) all ^= ev[w];
    asm volatile("" : : "s"(all));
    if (c.state != 0 || q == 0) return 0;
    return bs_control_step(c, ev, (fault & FAULT_FORCE) != 0);
}

// Member j takes its state at attempt q-1's accepted level (all six loads
// issued before the stores).
__device__ inline void bs_loop_accept(const BsArgs& A, const BsLoop& B, int q, int level, int j) {
    const long long nc = A.ncap;
    const double* cp = B.cand + (long long)((q + 1) & 1) * 54 * nc + (long long)(level - 4) * 6 * nc;
    double st[6];
    for (int qc = 0; qc < 6; qc++) st[qc] = cp[qc * nc + j];
    for (int qc = 0; qc < 3; qc++) {
        A.x[qc * nc + j] = st[qc];
        A.v[qc * nc + j] = st[3 + qc];
    }
}

// Attempt q's step, candidate set and error set.
__device__ inline BsArgs bs_loop_attempt(const BsArgs& A, const BsLoop& B, int q, const BsCtl& c) {
    BsArgs a = A;
    a.h0 = c.hlocal;
    a.cand = B.cand + (long long)(q & 1) * 54 * A.ncap;
    a.err = B.err + 12 * (q % 3);
    return a;
}

__global__ void __launch_bounds__(kLevels * kEncBig) bs2_loop_big(BsArgs A, BsLoop B, int q) {
    BsCtl c;
    const int level = bs_loop_decide(A, B, q, c);
    const int t = threadIdx.x;
    if (t == 0) B.ctl[(q + 1) & 1] = c;
    if (t < 12) B.err[12 * ((q + 1) % 3) + t] = 0ull;
    if (level >= 4 && t < A.nbig - 1) bs_loop_accept(A, B, q, level, 1 + t);
    if (c.state != 0) return;
    // the attempt starts from the accepted candidates (being stored into x, v
    // above for bs2_loop_particles) or from the unchanged state
    const long long nc = A.ncap;
    const double* cp =
        level >= 4 ? B.cand + (long long)((q + 1) & 1) * 54 * nc + (long long)(level - 4) * 6 * nc : nullptr;
    bs2_big_body(bs_loop_attempt(A, B, q, c), level >= 4 ? cp : A.x, level >= 4 ? cp + 3 * nc : A.v);
}

__global__ void __launch_bounds__(kBsPairThreads) bs2_loop_particles(BsArgs A, BsLoop B, int q) {
    const int t = threadIdx.x;
    const int p = blockIdx.x * kBsPairSlots + t % kBsPairSlots;
    const RegPartners RP = p < A.n - A.nbig ? load_partners(A, p) : RegPartners{0, 1, 1, 0.0, 0.0, 0.0, 0.0};
    BsCtl c;
    const int level = bs_loop_decide(A, B, q, c);
    if (level >= 4 && t < kBsPairSlots && p < A.n - A.nbig) bs_loop_accept(A, B, q, level, A.nbig + p);
    if (c.state != 0) return;
    // the attempt starts from the accepted candidates (those just stored into
    // x, v by the threads above) or from the unchanged state
    const long long nc = A.ncap;
    const double* cp =
        level >= 4 ? B.cand + (long long)((q + 1) & 1) * 54 * nc + (long long)(level - 4) * 6 * nc : nullptr;
    bs2_particles_body(bs_loop_attempt(A, B, q, c), level >= 4 ? cp : A.x, level >= 4 ? cp + 3 * nc : A.v, RP);
}

// The loop in one launch per attempt with bs2_particles' level pairing: every
// block first advances the Big members' 12 levels in LDS (thread (level,
// member), one force evaluation per round) and keeps their positions at all
// 78 force points (TB, dynamic LDS of 78 x 3 x nb doubles); then its
// kBsPairSlots particles run their paired levels from that table.  Block 0
// alone applies the Big members' accept, extrapolates them and stores the
// state.  Results are bs2_big's and bs2_particles', bit for bit.
__global__ void __launch_bounds__(kBsPairThreads) bs2_loop_paired(BsArgs A, BsLoop B, int q) {
    constexpr int NS = kBsPairSlots;
    static_assert(kLevels * kEncBig == kBsPairThreads, "thread (level, Big member) in the Big rounds");
    extern __shared__ double TB[];
    const int t = threadIdx.x;
    const int nb = A.nbig - 1;
    const int np = A.n - A.nbig;
    const bool lead = blockIdx.x == 0;
    const long long nc = A.ncap;
    const int sl = t % NS, pair = t / NS;
    const int pi_ = blockIdx.x * NS + sl;
    const bool live = pi_ < np;
    const int me = A.nbig + pi_;
    const RegPartners RP = live ? load_partners(A, pi_) : RegPartners{0, 1, 1, 0.0, 0.0, 0.0, 0.0};
    BsCtl c;
    const int level = bs_loop_decide(A, B, q, c);
    if (lead && t == 0) B.ctl[(q + 1) & 1] = c;
    if (lead && t < 12) B.err[12 * ((q + 1) % 3) + t] = 0ull;
    if (level >= 4) {
        if (pair == 0 && live) bs_loop_accept(A, B, q, level, me);
        if (pair == 1 && lead && sl < nb) bs_loop_accept(A, B, q, level, 1 + sl);
    }
    if (c.state != 0) return;
    const BsArgs a = bs_loop_attempt(A, B, q, c);
    const double h0 = a.h0;
    // the attempt starts from the accepted candidates or from the unchanged state
    const double* cp =
        level >= 4 ? B.cand + (long long)((q + 1) & 1) * 54 * nc + (long long)(level - 4) * 6 * nc : nullptr;
    const double* xs = level >= 4 ? cp : A.x;
    const double* vs = level >= 4 ? cp + 3 * nc : A.v;
    __shared__ double X0[3 * kEncBig];
    __shared__ double D[72 * NS];            // [6 * level + component][slot], then [..][Big member]
    static_assert(6 * 9 * NS <= 72 * NS, "the level errors' d(1) table fits in D");
    __shared__ double XS[NS][2];
    __shared__ double BXS[kEncBig][2];
    __shared__ unsigned long long EM[9];
    __shared__ BigShared S;
    __shared__ double W[kExPairs][2];
    if (t < 9) EM[t] = 0ull;
    stage_big(A, S, t, kBsPairThreads);
    ex_weights(h0, W, t, kBsPairThreads);
    // Big member k+1 at level n
    const int n = t / kEncBig + 1, k = t % kEncBig;
    const bool mem = k < nb;
    double bx0[3] = {0, 0, 0}, bv0[3] = {0, 0, 0}, ba0[3] = {0, 0, 0};
    if (mem) {
        for (int qc = 0; qc < 3; qc++) {
            bx0[qc] = xs[qc * nc + 1 + k];
            bv0[qc] = vs[qc * nc + 1 + k];
            if (n == 1) X0[3 * k + qc] = bx0[qc];
        }
        if (n == 1) bs_scales(bx0, bv0, BXS[k][0], BXS[k][1]);
    }
    double x0[3] = {1, 0, 0}, v0[3] = {0, 1, 0}, a0[3] = {0, 0, 0};
    double rcj = 0.0;
    int p0 = 0, p1 = 0;
    if (live) {
        for (int qc = 0; qc < 3; qc++) { x0[qc] = xs[qc * nc + me]; v0[qc] = vs[qc * nc + me]; }
        if (RP.n < 0) {
            rcj = A.rcrit[me];
            p0 = A.poff[pi_];
            p1 = A.poff[pi_ + 1];
        }
        if (pair == 0) bs_scales(x0, v0, XS[sl][0], XS[sl][1]);
    }
    const double m0 = A.m[0];
    __syncthreads();
    if (mem) big_accel(S, A.nbbp, k + 1, X0, ba0);
    BsLevel LB;
    bs_level_start(LB, n, h0, bx0, bv0, ba0);
    double ab[3] = {0, 0, 0};
    for (int r = 1; r <= kLevels; r++) {
        const bool act = mem && r <= n;
        // every (level, round) has its own point in TB, so a round's writes never
        // meet the previous round's reads: one barrier per round
        double* Pt = TB + bs_point(n, r) * 3 * nb;
        if (act)
            for (int qc = 0; qc < 3; qc++) Pt[3 * k + qc] = LB.xend[qc];
        __syncthreads();
        if (act) big_accel(S, A.nbbp, k + 1, Pt, ab);
        if (act && r < n) bs_level_step(LB, r, ab, bx0, bv0, ba0);
    }
    // particles: mfo_hkce with the Big positions Xb(i)
    auto pforce = [&](const double* xe, double* acc_out, auto Xb) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int qq = p0; qq < p1; qq++) {
            int i = A.pidx[qq];
            const double* Xi = Xb(i);
            double dx = xe[0] - Xi[0], dy = xe[1] - Xi[1], dz = xe[2] - Xi[2];
            double tmp2;
            if (!hkce_factor(dx * dx + dy * dy + dz * dz, fmax(A.rcrit[i], rcj), tmp2)) continue;
            double faci = tmp2 * A.m[i];
            acc[0] = acc[0] - faci * dx; acc[1] = acc[1] - faci * dy; acc[2] = acc[2] - faci * dz;
        }
        double s2 = xe[0] * xe[0] + xe[1] * xe[1] + xe[2] * xe[2];
        double s_1 = rsqrt_exact(s2);
        double tmp2 = A.m[0] * s_1 * s_1 * s_1;
        for (int qc = 0; qc < 3; qc++) acc_out[qc] = acc[qc] - tmp2 * xe[qc];
    };
    if (live) {
        if (RP.n >= 0) pforce_reg(RP, x0, X0 + 3 * (RP.i0 - 1), X0 + 3 * (RP.i1 - 1), m0, a0);
        else pforce(x0, a0, [&](int i) { return (const double*)(X0 + 3 * (i - 1)); });
#pragma unroll 1
        for (int kk = 0; kk < 2; kk++) {
            const int nn = kk == 0 ? pair + 1 : kLevels - pair;
            BsLevel L;
            bs_level_start(L, nn, h0, x0, v0, a0);
            double ap[3];
            for (int r = 1; r <= nn; r++) {
                const double* T = TB + bs_point(nn, r) * 3 * nb;
                if (RP.n >= 0) pforce_reg(RP, L.xend, T + 3 * (RP.i0 - 1), T + 3 * (RP.i1 - 1), m0, ap);
                else pforce(L.xend, ap, [&](int i) { return T + 3 * (i - 1); });
                if (r < nn) bs_level_step(L, r, ap, x0, v0, a0);
            }
            double R[6];
            bs_level_result(L, ap, v0, R);
            for (int qc = 0; qc < 6; qc++) D[(6 * (nn - 1) + qc) * NS + sl] = R[qc];
        }
    }
    __syncthreads();
    // particle extrapolation: thread (slot, component); then d(1) of levels
    // 4..12 go into D's space, no longer read, for the level errors
    double* E = D;                           // [slot or Big member][component][level]
    {
        const int s = t / 6, qc = t % 6;
        const int ps = blockIdx.x * NS + s;
        double e[9];
        if (ps < np) bs_extrapolate(W, D + qc * NS + s, NS, e, a.cand + (long long)qc * nc + A.nbig + ps, nc);
        __syncthreads();
        if (ps < np)
            for (int l = 0; l < 9; l++) E[(6 * s + qc) * 9 + l] = e[l];
    }
    __syncthreads();
    // level errors: item (slot, level 4 + l), merged per level over the block
    for (int w = t; w < 9 * NS; w += kBsPairThreads) {
        const int s = w / 9, l = w % 9;
        const int ps = blockIdx.x * NS + s;
        if (ps < np) {
            double d1[6];
            for (int qc = 0; qc < 6; qc++) d1[qc] = E[(6 * s + qc) * 9 + l];
            double err = bs_level_err(d1, XS[s][0], XS[s][1], A.fault);
            atomicMax(&EM[l], (unsigned long long)__double_as_longlong(err));
        }
    }
    __syncthreads();
    if (t < 9 && EM[t]) atomicMax(a.err + 3 + t, EM[t]);
    if (!lead) return;
    // block 0: the Big members' d rows, extrapolation and level errors
    if (mem) {
        double R[6];
        bs_level_result(LB, ab, bv0, R);
        for (int qc = 0; qc < 6; qc++) D[(6 * (n - 1) + qc) * kEncBig + k] = R[qc];
    }
    __syncthreads();
    {
        const int km = t / 6, qc = t % 6;
        const bool valid = t < 6 * kEncBig && km < nb;
        double e[9];
        if (valid) bs_extrapolate(W, D + qc * kEncBig + km, kEncBig, e, a.cand + (long long)qc * nc + 1 + km, nc);
        __syncthreads();
        if (valid)
            for (int l = 0; l < 9; l++) E[(6 * km + qc) * 9 + l] = e[l];
    }
    __syncthreads();
    if (t < 9 * kEncBig) {
        const int km = t / 9, l = t % 9;
        if (km < nb) {
            double d1[6];
            for (int qc = 0; qc < 6; qc++) d1[qc] = E[(6 * km + qc) * 9 + l];
            double err = bs_level_err(d1, BXS[km][0], BXS[km][1], A.fault);
            atomicMax(a.err + 3 + l, (unsigned long long)__double_as_longlong(err));
        }
    }
}

// The same loop in one launch per attempt, for encounter sets whose particle
// blocks all fit on the device at once (bs2_loop_fused): every block advances
// the Big members' 12 levels itself, in LDS, round by round beside its
// particles -- thread (level n, slot k) is Big member k+1 and particle slot k
// -- so no launch waits on another for the Big positions.  Block 0 alone
// applies the Big members' accept, extrapolates them and stores the state;
// the other blocks read the Big members' accepted states from the candidates
// (never from x, v, which block 0 may be writing).  Results are bs2_big's and
// bs2_particles', bit for bit.
__global__ void __launch_bounds__(kLevels * kBsPart) bs2_loop_fused(BsArgs A, BsLoop B, int q) {
    static_assert(kBsPart == kEncBig, "slot k is both Big member k+1 and particle slot k");
    const int t = threadIdx.x;
    const int n = t / kBsPart + 1;
    const int k = t % kBsPart;
    const int nb = A.nbig - 1;
    const int np = A.n - A.nbig;
    const bool lead = blockIdx.x == 0;
    const long long nc = A.ncap;
    BsCtl c;
    const int level = bs_loop_decide(A, B, q, c);
    if (lead && t == 0) B.ctl[(q + 1) & 1] = c;
    if (lead && t < 12) B.err[12 * ((q + 1) % 3) + t] = 0ull;
    // the start of attempt q: attempt q-1's accepted states (level >= 4) or the unchanged ones
    const double* cp =
        level >= 4 ? B.cand + (long long)((q + 1) & 1) * 54 * nc + (long long)(level - 4) * 6 * nc : nullptr;
    auto start = [&](int j, int comp) {
        return level >= 4 ? cp[comp * nc + j] : (comp < 3 ? A.x[comp * nc + j] : A.v[(comp - 3) * nc + j]);
    };
    const int pi_ = blockIdx.x * kBsPart + k;
    const bool live = pi_ < np;
    const int me = A.nbig + pi_;
    const bool mem = k < nb;
    if (level >= 4) {
        if (n == 1 && live) bs_loop_accept(A, B, q, level, me);
        if (n == 2 && lead && mem) bs_loop_accept(A, B, q, level, 1 + k);
    }
    if (c.state != 0) return;
    const BsArgs a = bs_loop_attempt(A, B, q, c);
    const double h0 = a.h0;
    __shared__ double P[kLevels][3 * kEncBig];
    __shared__ double X0[3 * kEncBig];
    __shared__ double D[72 * kBsPart];       // [6 * level + component][slot]
    __shared__ double E[kBsPart][6][9];
    __shared__ unsigned long long EM[9];
    __shared__ BigShared S;
    __shared__ double W[kExPairs][2];
    if (t < 9) EM[t] = 0ull;
    stage_big(A, S, t, kLevels * kBsPart);
    ex_weights(h0, W, t, kLevels * kBsPart);
    double bx0[3] = {0, 0, 0}, bv0[3] = {0, 0, 0}, ba0[3] = {0, 0, 0};
    if (mem)
        for (int qc = 0; qc < 3; qc++) {
            bx0[qc] = start(1 + k, qc);
            bv0[qc] = start(1 + k, 3 + qc);
            if (n == 1) X0[3 * k + qc] = bx0[qc];
        }
    double x0[3] = {1, 0, 0}, v0[3] = {0, 1, 0}, a0[3] = {0, 0, 0};
    double rcj = 0.0;
    int p0 = 0, p1 = 0;
    const RegPartners RP = live ? load_partners(A, pi_) : RegPartners{0, 1, 1, 0.0, 0.0, 0.0, 0.0};
    if (live) {
        for (int qc = 0; qc < 3; qc++) { x0[qc] = start(me, qc); v0[qc] = start(me, 3 + qc); }
        if (RP.n < 0) {
            rcj = A.rcrit[me];
            p0 = A.poff[pi_];
            p1 = A.poff[pi_ + 1];
        }
    }
    // mfo_hkce for a particle (as bs2_particles); Xb holds the Big members' positions
    auto pforce = [&](const double* xe, double* acc_out, const double* Xb) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int qq = p0; qq < p1; qq++) {
            int i = A.pidx[qq];
            const double* Xi = Xb + 3 * (i - 1);
            double dx = xe[0] - Xi[0], dy = xe[1] - Xi[1], dz = xe[2] - Xi[2];
            double tmp2;
            if (!hkce_factor(dx * dx + dy * dy + dz * dz, fmax(A.rcrit[i], rcj), tmp2)) continue;
            double faci = tmp2 * A.m[i];
            acc[0] = acc[0] - faci * dx; acc[1] = acc[1] - faci * dy; acc[2] = acc[2] - faci * dz;
        }
        double s2 = xe[0] * xe[0] + xe[1] * xe[1] + xe[2] * xe[2];
        double s_1 = rsqrt_exact(s2);
        double tmp2 = A.m[0] * s_1 * s_1 * s_1;
        for (int qc = 0; qc < 3; qc++) acc_out[qc] = acc[qc] - tmp2 * xe[qc];
    };
    const double m0 = A.m[0];
    __syncthreads();
    if (mem) big_accel(S, A.nbbp, k + 1, X0, ba0);
    if (live) {
        if (RP.n >= 0) pforce_reg(RP, x0, X0 + 3 * (RP.i0 - 1), X0 + 3 * (RP.i1 - 1), m0, a0);
        else pforce(x0, a0, X0);
    }
    BsLevel LB, LP;
    bs_level_start(LB, n, h0, bx0, bv0, ba0);
    bs_level_start(LP, n, h0, x0, v0, a0);
    double ab[3] = {0, 0, 0}, ap[3] = {0, 0, 0};
    for (int r = 1; r <= kLevels; r++) {
        const bool act = r <= n;
        if (act && mem)
            for (int qc = 0; qc < 3; qc++) P[n - 1][3 * k + qc] = LB.xend[qc];
        __syncthreads();
        if (act && mem) big_accel(S, A.nbbp, k + 1, P[n - 1], ab);
        if (act && live) {
            if (RP.n >= 0) {
                const double* X0 = P[n - 1] + 3 * (RP.i0 - 1);
                const double* X1 = P[n - 1] + 3 * (RP.i1 - 1);
                pforce_reg(RP, LP.xend, X0, X1, m0, ap);
            } else {
                pforce(LP.xend, ap, P[n - 1]);
            }
        }
        __syncthreads();
        if (act && r < n) {
            if (mem) bs_level_step(LB, r, ab, bx0, bv0, ba0);
            if (live) bs_level_step(LP, r, ap, x0, v0, a0);
        }
    }
    // particles: d rows, extrapolation (thread (slot, component)), level errors
    if (live) {
        double R[6];
        bs_level_result(LP, ap, v0, R);
        for (int qc = 0; qc < 6; qc++) D[(6 * (n - 1) + qc) * kBsPart + k] = R[qc];
    }
    __syncthreads();
    if (t < 6 * kBsPart) {
        const int sl = t / 6, qc = t % 6;
        const int ps = blockIdx.x * kBsPart + sl;
        if (ps < np) {
            double e[9];
            bs_extrapolate(W, D + qc * kBsPart + sl, kBsPart, e, a.cand + (long long)qc * nc + A.nbig + ps, nc);
            for (int l = 0; l < 9; l++) E[sl][qc][l] = e[l];
        }
    }
    __syncthreads();
    if (t < 9 * kBsPart) {
        const int sl = t / 9, l = t % 9;
        const int ps = blockIdx.x * kBsPart + sl;
        if (ps < np) {
            const int mp = A.nbig + ps;
            double xs, vs, xm[3], vm[3];
            for (int qc = 0; qc < 3; qc++) { xm[qc] = start(mp, qc); vm[qc] = start(mp, 3 + qc); }
            bs_scales(xm, vm, xs, vs);
            double d1[6];
            for (int qc = 0; qc < 6; qc++) d1[qc] = E[sl][qc][l];
            double err = bs_level_err(d1, xs, vs, A.fault);
            atomicMax(&EM[l], (unsigned long long)__double_as_longlong(err));
        }
    }
    __syncthreads();
    if (t < 9 && EM[t]) atomicMax(a.err + 3 + t, EM[t]);
    if (!lead) return;
    // block 0: the Big members' d rows, extrapolation and level errors (as bs2_big)
    if (mem) {
        double R[6];
        bs_level_result(LB, ab, bv0, R);
        for (int qc = 0; qc < 6; qc++) D[(6 * (n - 1) + qc) * kEncBig + k] = R[qc];
    }
    __syncthreads();
    if (t < 6 * kEncBig) {
        const int km = t / 6, qc = t % 6;
        if (km < nb) {
            double e[9];
            bs_extrapolate(W, D + qc * kEncBig + km, kEncBig, e, a.cand + (long long)qc * nc + 1 + km, nc);
            for (int l = 0; l < 9; l++) E[km][qc][l] = e[l];
        }
    }
    __syncthreads();
    if (t < 9 * kEncBig) {
        const int km = t / 9, l = t % 9;
        if (km < nb) {
            double xs, vs, xm[3], vm[3];
            for (int qc = 0; qc < 3; qc++) { xm[qc] = start(1 + km, qc); vm[qc] = start(1 + km, 3 + qc); }
            bs_scales(xm, vm, xs, vs);
            double d1[6];
            for (int qc = 0; qc < 6; qc++) d1[qc] = E[km][qc][l];
            double err = bs_level_err(d1, xs, vs, A.fault);
            atomicMax(a.err + 3 + l, (unsigned long long)__double_as_longlong(err));
        }
    }
}

}  // namespace mhip
