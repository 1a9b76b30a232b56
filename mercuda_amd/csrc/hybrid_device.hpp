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
    if (p < np)
        for (int c = 0; c < 3; c++) T[kHyHdr + HB_N * p + HB_XS + c] = X[3 * p + c];
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
            V[3 * p + c] = V[3 * p + c] + hby2 * (0.0 + 0.0 + A[3 * p + c]);
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
    double x0[3], v0[3];
    for (int c = 0; c < 3; c++) {
        args.st.ox[c * cap + g] = x[c];
        args.st.ov[c * cap + g] = v[c] + T[c];
        if (args.reset) args.st.a[c * cap + g] = a[c];
        v[c] = v[c] + hby2 * (0.0 + 0.0 + a[c]);
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
    if (p < np)
        for (int c = 0; c < 3; c++) V[3 * p + c] = V[3 * p + c] + hby2 * (0.0 + 0.0 + A[3 * p + c]);
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
    bool bad = false;
    for (int c = 0; c < 3; c++) {
        double v = args.st.v[c * cap + g] + d.hby2 * (0.0 + 0.0 + a[c]);
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
    double h0, tol2;
    double* dbuf;               // per particle d table [72][ncap]
    double* cand;               // accepted states per level n = 4..12: [9][6][ncap]
    unsigned long long* err;    // max error per level [12] (bits of non-negative doubles)
    int* fault;
};

// mfo_hkce pair factor (1 - K) / s^3 inside rc, else 0 (mercury6_2.for, mfo_hkce).
__device__ __forceinline__ bool hkce_factor(double s2, double rc, double& tmp2) {
    double rc2 = rc * rc;
    if (!(s2 < rc2)) return false;
    double s_1 = 1.0 / sqrt(s2);
    double s_3 = s_1 * s_1 * s_1;
    if (s2 <= 0.01 * rc2) {
        tmp2 = s_3;
    } else {
        double s = 1.0 / s_1;
        double q = (s - 0.1 * rc) / (0.9 * rc);
        double q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3;
        tmp2 = (1.0 - 10.0 * q3 + 15.0 * q4 - 6.0 * q5) * s_3;
    }
    return true;
}

__device__ __forceinline__ void merge_err(unsigned long long* slot, double e) {
    unsigned long long bits = (unsigned long long)__double_as_longlong(e);
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(bits, off);
        bits = o > bits ? o : bits;
    }
    if (__lane_id() == 0) atomicMax(slot, bits);
}

// Big members' forces at positions P (LDS, [3][kEncBig]), lane k = member k+1.
__device__ inline void bs_big_force(const BsArgs& a, int nb, const double* P, double* F) {
    const int k = threadIdx.x;
    if (k < nb) {
        double acc[3] = {0.0, 0.0, 0.0};
        const int me = k + 1;
        for (int q = 0; q < a.nbbp; q++) {
            int i = a.bbp[2 * q], j = a.bbp[2 * q + 1];
            if (i != me && j != me) continue;
            double dx = P[3 * (j - 1)] - P[3 * (i - 1)];
            double dy = P[3 * (j - 1) + 1] - P[3 * (i - 1) + 1];
            double dz = P[3 * (j - 1) + 2] - P[3 * (i - 1) + 2];
            double tmp2;
            if (!hkce_factor(dx * dx + dy * dy + dz * dz, fmax(a.rcrit[i], a.rcrit[j]), tmp2)) continue;
            if (j == me) {
                double faci = tmp2 * a.m[i];
                acc[0] = acc[0] - faci * dx; acc[1] = acc[1] - faci * dy; acc[2] = acc[2] - faci * dz;
            } else {
                double facj = tmp2 * a.m[j];
                acc[0] = acc[0] + facj * dx; acc[1] = acc[1] + facj * dy; acc[2] = acc[2] + facj * dz;
            }
        }
        const double* x = P + 3 * k;
        double s2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        double s_1 = 1.0 / sqrt(s2);
        double tmp2 = a.m[0] * s_1 * s_1 * s_1;
        for (int c = 0; c < 3; c++) F[3 * k + c] = acc[c] - tmp2 * x[c];
    }
}

// One member's 12-level modified-midpoint sequence with polynomial
// extrapolation (mdt_bs2 mercury6_2.for:3470-3570): per level n > 3 its
// error and accepted state.  force(point, xend, a) gives the acceleration at
// force point `point`; D is this member's d table (6 x 12, strided).
template <class Force>
__device__ inline void bs_member(const BsArgs& A, double h0, const double* x0, const double* v0, const double* a0,
                                 Force force, double* D, long long ds, double xscal, double vscal, double* errs,
                                 double* cands, long long cs) {
    double h2[12];
    int point = 0;
    for (int n = 1; n <= 12; n++) {
        double h = h0 / (double)n;
        double hby2 = .5 * h;
        h2[n - 1] = h * h;
        double h2by2 = .5 * h2[n - 1];
        double b[3], c[3], xend[3], a[3];
        for (int q = 0; q < 3; q++) {
            b[q] = .5 * a0[q];
            c[q] = 0.0;
            xend[q] = h2by2 * a0[q] + h * v0[q] + x0[q];
        }
        for (int j = 2; j <= n; j++) {
            force(point++, xend, a);
            double tmp0 = h * (double)j;
            for (int q = 0; q < 3; q++) {
                b[q] = b[q] + a[q];
                c[q] = c[q] + b[q];
                xend[q] = h2[n - 1] * c[q] + h2by2 * a0[q] + tmp0 * v0[q] + x0[q];
            }
        }
        force(point++, xend, a);
        for (int q = 0; q < 3; q++) {
            D[(long long)(6 * (n - 1) + q) * ds] = xend[q];
            D[(long long)(6 * (n - 1) + 3 + q) * ds] = h * b[q] + hby2 * a[q] + v0[q];
        }
        for (int j = n - 1; j >= 1; j--) {
            int j1 = j + 1;
            double tmp0 = 1.0 / (h2[j - 1] - h2[n - 1]);
            double tmp1 = tmp0 * h2[j1 - 1];
            double tmp2 = tmp0 * h2[n - 1];
            for (int q = 0; q < 6; q++)
                D[(long long)(6 * (j - 1) + q) * ds] = tmp1 * D[(long long)(6 * (j1 - 1) + q) * ds]
                                                     - tmp2 * D[(long long)(6 * (j - 1) + q) * ds];
        }
        if (n > 3) {
            double d1[6];
            for (int q = 0; q < 6; q++) d1[q] = D[(long long)q * ds];
            bool fin = true;
            for (int q = 0; q < 6; q++) fin = fin && isfinite(d1[q]);
            double tmp1 = fmax(fmax(d1[0] * d1[0], d1[1] * d1[1]), d1[2] * d1[2]);
            double tmp2 = fmax(fmax(d1[3] * d1[3], d1[4] * d1[4]), d1[5] * d1[5]);
            if (!fin) atomicOr(A.fault, FAULT_FORCE);
            errs[n - 4] = fin ? fmax(tmp1 * xscal, tmp2 * vscal) : 0.0;
            for (int q = 0; q < 6; q++) {
                double s = d1[q];
                for (int j = 2; j <= n; j++) s = s + D[(long long)(6 * (j - 1) + q) * ds];
                cands[(long long)((n - 4) * 6 + q) * cs] = s;
            }
        }
    }
}

__device__ __forceinline__ void bs_scales(const double* x0, const double* v0, double& xs, double& vs) {
    double t1 = x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2];
    double t2 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
    xs = 1.0 / fmax(t1, 1e-300);
    vs = 1.0 / fmax(t2, 1e-300);
}

// One BS2 attempt of the whole encounter set.  Every workgroup first runs the
// Big members' sequence (one lane each, identical in every workgroup) into an
// LDS table of their positions at the 78 force points; then each lane runs one
// particle member's sequence against that table.
__global__ void __launch_bounds__(64) bs2_attempt(BsArgs A) {
    const int nb = A.nbig - 1;
    const int lane = threadIdx.x;
    __shared__ double TP[kBsPoints][3 * kEncBig];
    __shared__ double P[3 * kEncBig], F[3 * kEncBig];
    // d tables and accepted states of the Big members; column kEncBig is
    // scratch for the lanes past the Big members
    constexpr int kCol = kEncBig + 1;
    __shared__ double BD[72 * kCol];
    __shared__ double BC[54 * kCol];
    // --- Big members: every force point is a lockstep of all Big lanes
    {
        const int k = lane;
        double x0[3] = {0, 0, 0}, v0[3] = {0, 0, 0}, a0[3] = {0, 0, 0};
        if (k < nb)
            for (int c = 0; c < 3; c++) {
                x0[c] = A.x[c * A.ncap + 1 + k];
                v0[c] = A.v[c * A.ncap + 1 + k];
                P[3 * k + c] = x0[c];
            }
        __syncthreads();
        bs_big_force(A, nb, P, F);
        __syncthreads();
        if (k < nb)
            for (int c = 0; c < 3; c++) a0[c] = F[3 * k + c];
        __syncthreads();
        double xs = 0, vs = 0;
        if (k < nb) bs_scales(x0, v0, xs, vs);
        double errs[9];
        // the sequence with a force that synchronises the wave at every point
        auto force = [&](int point, const double* xe, double* a) {
            __syncthreads();
            if (k < nb)
                for (int c = 0; c < 3; c++) { P[3 * k + c] = xe[c]; TP[point][3 * k + c] = xe[c]; }
            __syncthreads();
            bs_big_force(A, nb, P, F);
            __syncthreads();
            for (int c = 0; c < 3; c++) a[c] = k < nb ? F[3 * k + c] : 0.0;
        };
        // lanes past nb run the same control flow on dummy data
        double dummy0[3] = {1, 0, 0}, dummyv[3] = {0, 1, 0};
        const double* mx0 = k < nb ? x0 : dummy0;
        const double* mv0 = k < nb ? v0 : dummyv;
        const int col = k < nb ? k : kEncBig;
        bs_member(A, A.h0, mx0, mv0, a0, force, BD + col, kCol, xs, vs, errs, BC + col, kCol);
        if (k < nb) {
            for (int q = 0; q < 9; q++) atomicMax(A.err + 3 + q, (unsigned long long)__double_as_longlong(errs[q]));
            if (blockIdx.x == 0)
                for (int q = 0; q < 54; q++) A.cand[(long long)q * A.ncap + 1 + k] = BC[q * kCol + k];
        }
        __syncthreads();
    }
    // --- particle members
    const int np = A.n - A.nbig;
    const int pi_ = blockIdx.x * 64 + lane;
    const bool live = pi_ < np;
    const int me = A.nbig + pi_;
    double x0[3] = {1, 0, 0}, v0[3] = {0, 1, 0}, a0[3];
    double rcj = 0.0;
    int p0 = 0, p1 = 0;
    if (live) {
        for (int c = 0; c < 3; c++) { x0[c] = A.x[c * A.ncap + me]; v0[c] = A.v[c * A.ncap + me]; }
        rcj = A.rcrit[me];
        p0 = A.poff[pi_];
        p1 = A.poff[pi_ + 1];
    }
    auto pforce = [&](int point, const double* xe, double* a) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int q = p0; q < p1; q++) {
            int i = A.pidx[q];
            const double* Xi = point >= 0 ? &TP[point][3 * (i - 1)] : &P[3 * (i - 1)];
            double dx = xe[0] - Xi[0], dy = xe[1] - Xi[1], dz = xe[2] - Xi[2];
            double tmp2;
            if (!hkce_factor(dx * dx + dy * dy + dz * dz, fmax(A.rcrit[i], rcj), tmp2)) continue;
            double faci = tmp2 * A.m[i];
            acc[0] = acc[0] - faci * dx; acc[1] = acc[1] - faci * dy; acc[2] = acc[2] - faci * dz;
        }
        double s2 = xe[0] * xe[0] + xe[1] * xe[1] + xe[2] * xe[2];
        double s_1 = 1.0 / sqrt(s2);
        double tmp2 = A.m[0] * s_1 * s_1 * s_1;
        for (int c = 0; c < 3; c++) a[c] = acc[c] - tmp2 * xe[c];
    };
    // start positions of the Big members for a0
    if (lane < nb)
        for (int c = 0; c < 3; c++) P[3 * lane + c] = A.x[c * A.ncap + 1 + lane];
    __syncthreads();
    if (live) pforce(-1, x0, a0);
    double errs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (live) {
        double xs, vs;
        bs_scales(x0, v0, xs, vs);
        bs_member(A, A.h0, x0, v0, a0, pforce, A.dbuf + pi_, A.ncap, xs, vs, errs, A.cand + me, A.ncap);
    }
    for (int q = 0; q < 9; q++) merge_err(A.err + 3 + q, errs[q]);
}

__global__ void bs2_accept(BsArgs A, int level) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (k >= A.n) return;
    for (int q = 0; q < 3; q++) {
        A.x[q * A.ncap + k] = A.cand[(long long)((level - 4) * 6 + q) * A.ncap + k];
        A.v[q * A.ncap + k] = A.cand[(long long)((level - 4) * 6 + 3 + q) * A.ncap + k];
    }
}

}  // namespace mhip
