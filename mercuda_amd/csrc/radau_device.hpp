// RA15 (Everhart's RADAU) for massless particles on gfx950, FP64.
//
// One RA15 sequence (mdt_ra15, mercury6_2.for:4133-4500) with the force of
// mfo_all (mfo_grav with the optional central 1PN term, plus oblateness and
// the non-gravitational model; mercury6_2.for:4506-4707, 5106-5357).  Massless particles do not act on the
// Big bodies, so the planets' Gauss-Radau predictors at every (iteration,
// substep) are independent of the particles:
//   ra_planets   one wave: the planets' whole iteration; per evaluation point
//                it tabulates the planets' predicted positions, masses, the
//                indirect term sum_i -m_i x_i / r_i^3 and the oblateness
//                reaction, and seeds the global max |b7| with the planets';
//   ra_particles one lane per particle: the particle's iteration from the
//                table, its b updated in place, max |b7| merged by atomicMax;
//   the step-size law runs on the host (libm pow, as the CPU path);
//   ra_finish    every body: x, v at the end of the sequence and the
//                predicted b, e of the next one.
// Per component the arithmetic is the Fortran's, in its order.
#pragma once
#include <hip/hip_runtime.h>
#include "mvs_device.hpp"

namespace mhip {

constexpr int kRaEval = 1 + 6 * 7;          // start + up to six iterations of seven substeps

struct RaConst {
    double h[8], xc[8], vc[7], c[21], d[21], r[28];
};

// The Gauss-Radau spacings h and the series constants xc, vc, c, d, r of
// mdt_ra15's first call (mercury6_2.for:4160-4199), evaluated at compile time
// with the same double operations in the same order.  The kernels read them
// as literals: passed in the kernel arguments, the 93 doubles stayed in SGPRs
// that spilled into VGPR lanes, and every use became a v_readlane.
constexpr RaConst make_ra_constants() {
    constexpr double h[8] = {0.0, .0562625605369221, .1802406917368924, .3526247171131696,
                             .5471536263305554, .7342101772154105, .8853209468390958, .9775206135612875};
    constexpr double xc[8] = {.5, .1666666666666667, .08333333333333333, .05, .03333333333333333,
                              .02380952380952381, .01785714285714286, .01388888888888889};
    constexpr double vc[7] = {.5, .3333333333333333, .25, .2, .1666666666666667, .1428571428571429, .125};
    RaConst k{};
    for (int i = 0; i < 8; i++) { k.h[i] = h[i]; k.xc[i] = xc[i]; }
    for (int i = 0; i < 7; i++) k.vc[i] = vc[i];
    int n = 0;
    for (int j = 1; j < 8; j++)
        for (int q = 0; q < j; q++) k.r[n++] = 1.0 / (h[j] - h[q]);
    // 1-based Fortran indices: c(n), d(n) stored at [n - 1]
    k.c[0] = -h[1];
    k.d[0] = h[1];
    n = 1;
    for (int j = 3; j <= 7; j++) {
        n++;
        k.c[n - 1] = -h[j - 1] * k.c[n - j + 1];
        k.d[n - 1] = h[1] * k.d[n - j + 1];
        for (int q = 3; q <= j - 1; q++) {
            n++;
            k.c[n - 1] = k.c[n - j] - h[j - 1] * k.c[n - j + 1];
            k.d[n - 1] = k.d[n - j] + h[q - 1] * k.d[n - j + 1];
        }
        n++;
        k.c[n - 1] = k.c[n - j] - h[j - 1];
        k.d[n - 1] = k.d[n - j] + h[j - 1];
    }
    return k;
}

// RA15 state, component-major: b[(3*i + comp) ... ] laid out as [21][cap]
// with index (comp * 7 + q) * cap + body, likewise e; a1 [3][cap].
struct RaState {
    double* b;
    double* e;
    double* a1;
};

struct RaArgs {
    const Sys* sys;
    State st;
    RaState ra;
    double* tab;              // kRaEval entries of stride Sys::ts (header + force block)
    unsigned long long* bmax; // max |b7| over all bodies (bit pattern of a non-negative double)
    RaConst k;
    double t;                 // sequence size
    int niter;                // 6 after a reset, 2 otherwise; 0: forces only
    int reset;                // zero b (and e) first
    int pn;                   // central 1PN term, velocity predictors
    double light2;            // (speed of light in AU/day)^2
    // The previous sequence's finish (mercury6_2.for:4426-4489), deferred to
    // this launch when nothing read the state in between (pend): the start
    // state advanced over ptdid with its final b and start acceleration, and
    // the predicted e, b of this sequence with pq = this size / ptdid (no b - e
    // correction after a reset sequence, preset; none at all when this one
    // resets).  The start state is written back; b and e as usual.
    double ptdid = 0.0, pq = 0.0;
    int pend = 0, preset = 0;
};

// The finish of one component of one body (ra_finish's arithmetic): x1, v1
// advanced over tdid from the final b and the start acceleration a1; then,
// unless the next sequence resets, e and b predicted for it in place.
__device__ __forceinline__ void ra_finish_comp(double tdid, double q, bool preset, bool reset,
                                               double* b, double* e, double a1, double& x1, double& v1) {
    constexpr RaConst K = make_ra_constants();
    const double temp = tdid * tdid;
    const double* xc = K.xc;
    const double* vc = K.vc;
    x1 = (xc[7] * b[6] + xc[6] * b[5] + xc[5] * b[4] + xc[4] * b[3] + xc[3] * b[2] + xc[2] * b[1] + xc[1] * b[0]
          + xc[0] * a1) * temp + v1 * tdid + x1;
    v1 = (vc[6] * b[6] + vc[5] * b[5] + vc[4] * b[4] + vc[3] * b[3] + vc[2] * b[2] + vc[1] * b[1] + vc[0] * b[0]
          + a1) * tdid + v1;
    if (reset) {
        for (int i = 0; i < 7; i++) { b[i] = 0.0; e[i] = 0.0; }
        return;
    }
    const double q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q3 * q4;
    double s[7];
    for (int i = 0; i < 7; i++) s[i] = preset ? 0.0 : b[i] - e[i];
    e[0] = q * (b[6] * 7.0 + b[5] * 6.0 + b[4] * 5.0 + b[3] * 4.0 + b[2] * 3.0 + b[1] * 2.0 + b[0]);
    e[1] = q2 * (b[6] * 21.0 + b[5] * 15.0 + b[4] * 10.0 + b[3] * 6.0 + b[2] * 3.0 + b[1]);
    e[2] = q3 * (b[6] * 35.0 + b[5] * 20.0 + b[4] * 10.0 + b[3] * 4.0 + b[2]);
    e[3] = q4 * (b[6] * 35.0 + b[5] * 15.0 + b[4] * 5.0 + b[3]);
    e[4] = q5 * (b[6] * 21.0 + b[5] * 6.0 + b[4]);
    e[5] = q6 * (b[6] * 7.0 + b[5]);
    e[6] = q7 * b[6];
    for (int i = 0; i < 7; i++) b[i] = e[i] + s[i];
}

// Heliocentric acceleration of a massless body (mfo_grav for j > nbig, plus
// oblateness): direct terms from the Big bodies in order, then the indirect
// and central terms, then 1PN (pn) and oblateness (d.obl).  Ops: the particle
// kernel's RangeOps give the IEEE bits for every squared distance sq_ok
// admits and raise the force fault for the rest, where s^-3 (or s^2) has
// overflowed in the IEEE form as well.
// TP: const double* or ConstTab* (scalar loads of a table the wave shares).
template <bool PN, class Ops = IeeeOps, class TP>
__device__ __forceinline__ void grav_small(const Sys& d, TP T, int np, double light2,
                                           double x, double y, double z, double vx, double vy, double vz,
                                           double& ax, double& ay, double& az, int* fault) {
    const TP F = T + kHdr;
    ax = 0.0; ay = 0.0; az = 0.0;
    for (int p = 0; p < np; p++) {
        double dx = x - F[kF * p], dy = y - F[kF * p + 1], dz = z - F[kF * p + 2];
        double s2 = dx * dx + dy * dy + dz * dz;
        if (!Ops::sq_ok(s2)) atomicOr(fault, FAULT_FORCE);
        double s_1 = Ops::rcp(Ops::sqrt(s2));
        double s_3 = s_1 * s_1 * s_1;
        double tmp1 = s_3 * F[kF * p + 3];
        ax = ax - tmp1 * dx;
        ay = ay - tmp1 * dy;
        az = az - tmp1 * dz;
    }
    double s2 = x * x + y * y + z * z;
    if (!Ops::sq_ok(s2)) atomicOr(fault, FAULT_FORCE);
    double s_1 = Ops::rcp(Ops::sqrt(s2));
    double r3 = s_1 * s_1 * s_1;
    double tmp1 = d.mu * r3;
    if (!PN) {
        ax = ax + T[0] - tmp1 * x;
        ay = ay + T[1] - tmp1 * y;
        az = az + T[2] - tmp1 * z;
    } else {
        double rv = x * vx + y * vy + z * vz;
        double v2 = vx * vx + vy * vy + vz * vz;
        double f = tmp1 / light2;
        double radial = 4.0 * d.mu * s_1 - v2;
        ax = ax + T[0] - tmp1 * x + f * (radial * x + 4.0 * rv * vx);
        ay = ay + T[1] - tmp1 * y + f * (radial * y + 4.0 * rv * vy);
        az = az + T[2] - tmp1 * z + f * (radial * z + 4.0 * rv * vz);
    }
    if (d.obl) {
        double ox, oy, oz;
        obl_body<Ops>(d.mu, d.j2, d.j4, d.j6, x, y, z, ox, oy, oz);
        ax = ax + (ox - T[3]);
        ay = ay + (oy - T[4]);
        az = az + (oz - T[5]);
    }
}

// g from the b predicted at the last call (Everhart eq. 7; mercury6_2.for:4228-4240)
__device__ __forceinline__ void ra_g_from_b(const double* b, double* g) {
    constexpr RaConst K = make_ra_constants();
    const double* d = K.d;
    g[0] = b[6] * d[15] + b[5] * d[10] + b[4] * d[6] + b[3] * d[3] + b[2] * d[1] + b[1] * d[0] + b[0];
    g[1] = b[6] * d[16] + b[5] * d[11] + b[4] * d[7] + b[3] * d[4] + b[2] * d[2] + b[1];
    g[2] = b[6] * d[17] + b[5] * d[12] + b[4] * d[8] + b[3] * d[5] + b[2];
    g[3] = b[6] * d[18] + b[5] * d[13] + b[4] * d[9] + b[3];
    g[4] = b[6] * d[19] + b[5] * d[14] + b[4];
    g[5] = b[6] * d[20] + b[5];
    g[6] = b[6];
}

// Position and velocity predictor factors at substep h (Everhart eqs. 9, 10)
__device__ __forceinline__ void ra_sx(double t, double hj, double* s) {
    s[0] = t * hj;
    s[1] = s[0] * s[0] * .5;
    s[2] = s[1] * hj * .3333333333333333;
    s[3] = s[2] * hj * .5;
    s[4] = s[3] * hj * .6;
    s[5] = s[4] * hj * .6666666666666667;
    s[6] = s[5] * hj * .7142857142857143;
    s[7] = s[6] * hj * .75;
    s[8] = s[7] * hj * .7777777777777778;
}
__device__ __forceinline__ void ra_sv(double t, double hj, double* s) {
    s[0] = t * hj;
    s[1] = s[0] * hj * .5;
    s[2] = s[1] * hj * .6666666666666667;
    s[3] = s[2] * hj * .75;
    s[4] = s[3] * hj * .8;
    s[5] = s[4] * hj * .8333333333333333;
    s[6] = s[5] * hj * .8571428571428571;
    s[7] = s[6] * hj * .875;
}
__device__ __forceinline__ double ra_xpred(const double* s, const double* b, double a1, double v1, double x1) {
    return s[8] * b[6] + s[7] * b[5] + s[6] * b[4] + s[5] * b[3] + s[4] * b[2] + s[3] * b[1] + s[2] * b[0]
         + s[1] * a1 + s[0] * v1 + x1;
}
__device__ __forceinline__ double ra_vpred(const double* s, const double* b, double a1, double v1) {
    return s[7] * b[6] + s[6] * b[5] + s[5] * b[4] + s[4] * b[3] + s[3] * b[2] + s[2] * b[1] + s[1] * b[0]
         + s[0] * a1 + v1;
}

// g and b update after the force at substep j (Everhart eqs. 4, 5;
// mercury6_2.for:4295-4391).
__device__ __forceinline__ void ra_update(int j, double a, double a1, double* g, double* b) {
    constexpr RaConst K = make_ra_constants();
    const double* r = K.r;
    const double* c = K.c;
    double temp, gk = a - a1;
    switch (j) {
    case 2:
        temp = g[0];
        g[0] = (a - a1) * r[0];
        b[0] = b[0] + g[0] - temp;
        break;
    case 3:
        temp = g[1];
        g[1] = (gk * r[1] - g[0]) * r[2];
        temp = g[1] - temp;
        b[0] = b[0] + temp * c[0];
        b[1] = b[1] + temp;
        break;
    case 4:
        temp = g[2];
        g[2] = ((gk * r[3] - g[0]) * r[4] - g[1]) * r[5];
        temp = g[2] - temp;
        b[0] = b[0] + temp * c[1];
        b[1] = b[1] + temp * c[2];
        b[2] = b[2] + temp;
        break;
    case 5:
        temp = g[3];
        g[3] = (((gk * r[6] - g[0]) * r[7] - g[1]) * r[8] - g[2]) * r[9];
        temp = g[3] - temp;
        b[0] = b[0] + temp * c[3];
        b[1] = b[1] + temp * c[4];
        b[2] = b[2] + temp * c[5];
        b[3] = b[3] + temp;
        break;
    case 6:
        temp = g[4];
        g[4] = ((((gk * r[10] - g[0]) * r[11] - g[1]) * r[12] - g[2]) * r[13] - g[3]) * r[14];
        temp = g[4] - temp;
        b[0] = b[0] + temp * c[6];
        b[1] = b[1] + temp * c[7];
        b[2] = b[2] + temp * c[8];
        b[3] = b[3] + temp * c[9];
        b[4] = b[4] + temp;
        break;
    case 7:
        temp = g[5];
        g[5] = (((((gk * r[15] - g[0]) * r[16] - g[1]) * r[17] - g[2]) * r[18] - g[3]) * r[19] - g[4]) * r[20];
        temp = g[5] - temp;
        b[0] = b[0] + temp * c[10];
        b[1] = b[1] + temp * c[11];
        b[2] = b[2] + temp * c[12];
        b[3] = b[3] + temp * c[13];
        b[4] = b[4] + temp * c[14];
        b[5] = b[5] + temp;
        break;
    default:
        temp = g[6];
        g[6] = ((((((gk * r[21] - g[0]) * r[22] - g[1]) * r[23] - g[2]) * r[24] - g[3]) * r[25] - g[4]) * r[26]
                - g[5]) * r[27];
        temp = g[6] - temp;
        b[0] = b[0] + temp * c[15];
        b[1] = b[1] + temp * c[16];
        b[2] = b[2] + temp * c[17];
        b[3] = b[3] + temp * c[18];
        b[4] = b[4] + temp * c[19];
        b[5] = b[5] + temp * c[20];
        b[6] = b[6] + temp;
        break;
    }
}

__device__ __forceinline__ void merge_bmax(unsigned long long* bmax, double v, int* fault) {
    if (!isfinite(v)) atomicOr(fault, FAULT_FORCE);
    unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(bits, off);
        bits = o > bits ? o : bits;
    }
    if (__lane_id() == 0) atomicMax(bmax, bits);
}

// Pair terms of the planets' mutual forces computed once per pair, in
// parallel over the wave, for systems of up to kPairPlanets planets: s^-3 and
// the separation x_j - x_i (i < j), which is the dx of both bodies' terms in
// mfo_grav, so the sums below give the same bits as computing each in its lane.
constexpr int kPairPlanets = 16;
constexpr int kPairMax = kPairPlanets * (kPairPlanets - 1) / 2;

struct RaPairs {
    double* s3;            // [kPairMax]
    double* dx;            // [3][kPairMax]
    double* rc;            // [kMaxBig] 1/|x| of each planet (indirect and central terms)
    const unsigned char* pi;   // pair q = (pi[q], pj[q]), i < j, in row order
    const unsigned char* pj;
};

__device__ __forceinline__ int ra_pair_index(int np, int i, int j) { return i * np - i * (i + 1) / 2 + (j - i - 1); }

// Big-body accelerations of mfo_grav (mercury6_2.for:4616-4690) for the
// planets at X (LDS): lane p computes body p+1; lane 0 also writes the
// table header (indirect term, oblateness reaction) and the force block.
// With pair terms (np <= kPairPlanets) the separations, s^-3 and 1/|x| come
// from one parallel pass; each lane then sums its partners in the same order
// (i < p subtracted, then j > p added) with the same operations.
//
// NG: the Big bodies' non-gravitational terms (mfo_ngf, mfo_pr) are added
// last, as mfo_all adds them; ngf is the [5][cap] parameter array.
template <bool PN, bool NG = false>
__device__ void ra_planet_forces(const Sys& d, int np, const double* m, const double* X, const double* V,
                                 double* A, double* hdr, double* aob, double light2, double* T, int* fault,
                                 const RaPairs* P = nullptr, const double* ngf = nullptr, long long cap = 0) {
    const int lane = threadIdx.x;
    if (P) {
        const int npairs = np * (np - 1) / 2;
        for (int q = lane; q < npairs; q += blockDim.x) {
            const int i = P->pi[q], j = P->pj[q];
            double dx = X[3 * j] - X[3 * i], dy = X[3 * j + 1] - X[3 * i + 1], dz = X[3 * j + 2] - X[3 * i + 2];
            double s2 = dx * dx + dy * dy + dz * dz;
            if (!(s2 > 0.0)) atomicOr(fault, FAULT_FORCE);
            double s_1 = 1.0 / sqrt(s2);
            P->s3[q] = s_1 * s_1 * s_1;
            P->dx[q] = dx; P->dx[kPairMax + q] = dy; P->dx[2 * kPairMax + q] = dz;
        }
        if (lane < np) {
            const double* xb = X + 3 * lane;
            double s2 = xb[0] * xb[0] + xb[1] * xb[1] + xb[2] * xb[2];
            if (!(s2 > 0.0)) atomicOr(fault, FAULT_FORCE);
            P->rc[lane] = 1.0 / sqrt(s2);
        }
        __syncthreads();
        if (lane < np) {
            const double* xb = X + 3 * lane;
            double a[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < lane; i++) {
                const int q = ra_pair_index(np, i, lane);
                double tmp1 = P->s3[q] * m[i + 1];
                a[0] = a[0] - tmp1 * P->dx[q]; a[1] = a[1] - tmp1 * P->dx[kPairMax + q];
                a[2] = a[2] - tmp1 * P->dx[2 * kPairMax + q];
            }
            for (int j = lane + 1; j < np; j++) {
                const int q = ra_pair_index(np, lane, j);
                double tmp2 = P->s3[q] * m[j + 1];
                a[0] = a[0] + tmp2 * P->dx[q]; a[1] = a[1] + tmp2 * P->dx[kPairMax + q];
                a[2] = a[2] + tmp2 * P->dx[2 * kPairMax + q];
            }
            for (int c = 0; c < 3; c++) A[3 * lane + c] = a[c];
            if (d.obl) obl_body(d.mu, d.j2, d.j4, d.j6, xb[0], xb[1], xb[2], aob[3 * lane], aob[3 * lane + 1],
                                aob[3 * lane + 2]);
        }
    } else if (lane < np) {
        const double* xb = X + 3 * lane;
        double a[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < lane; i++) {
            double dx = xb[0] - X[3 * i], dy = xb[1] - X[3 * i + 1], dz = xb[2] - X[3 * i + 2];
            double s2 = dx * dx + dy * dy + dz * dz;
            if (!(s2 > 0.0)) atomicOr(fault, FAULT_FORCE);
            double s_1 = 1.0 / sqrt(s2);
            double s_3 = s_1 * s_1 * s_1;
            double tmp1 = s_3 * m[i + 1];
            a[0] = a[0] - tmp1 * dx; a[1] = a[1] - tmp1 * dy; a[2] = a[2] - tmp1 * dz;
        }
        for (int j = lane + 1; j < np; j++) {
            double dx = X[3 * j] - xb[0], dy = X[3 * j + 1] - xb[1], dz = X[3 * j + 2] - xb[2];
            double s2 = dx * dx + dy * dy + dz * dz;
            if (!(s2 > 0.0)) atomicOr(fault, FAULT_FORCE);
            double s_1 = 1.0 / sqrt(s2);
            double s_3 = s_1 * s_1 * s_1;
            double tmp2 = s_3 * m[j + 1];
            a[0] = a[0] + tmp2 * dx; a[1] = a[1] + tmp2 * dy; a[2] = a[2] + tmp2 * dz;
        }
        for (int c = 0; c < 3; c++) A[3 * lane + c] = a[c];
        if (d.obl) obl_body(d.mu, d.j2, d.j4, d.j6, xb[0], xb[1], xb[2], aob[3 * lane], aob[3 * lane + 1],
                            aob[3 * lane + 2]);
    }
    __syncthreads();
    if (lane == 0) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int i = 0; i < np; i++) {
            const double* xi = X + 3 * i;
            double s_1;
            if (P) {
                s_1 = P->rc[i];
            } else {
                double s2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
                if (!(s2 > 0.0)) atomicOr(fault, FAULT_FORCE);
                s_1 = 1.0 / sqrt(s2);
            }
            double tmp1 = m[i + 1] * (s_1 * s_1 * s_1);
            sx = sx - tmp1 * xi[0];
            sy = sy - tmp1 * xi[1];
            sz = sz - tmp1 * xi[2];
        }
        double acen[3] = {0.0, 0.0, 0.0};
        if (d.obl)
            for (int i = 0; i < np; i++) {
                double tmp4 = m[i + 1] / m[0];
                for (int c = 0; c < 3; c++) acen[c] = acen[c] - tmp4 * aob[3 * i + c];
            }
        hdr[0] = sx; hdr[1] = sy; hdr[2] = sz;
        hdr[3] = acen[0]; hdr[4] = acen[1]; hdr[5] = acen[2];
        for (int k = 0; k < 6; k++) T[k] = hdr[k];
        T[6] = 0.0; T[7] = 0.0;
    }
    __syncthreads();
    if (lane < np) {
        const double* xb = X + 3 * lane;
        const double* vb = V + 3 * lane;
        double s_1;
        if (P) {
            s_1 = P->rc[lane];
        } else {
            double s2 = xb[0] * xb[0] + xb[1] * xb[1] + xb[2] * xb[2];
            s_1 = 1.0 / sqrt(s2);
        }
        double r3 = s_1 * s_1 * s_1;
        double tmp1 = m[0] * r3;
        double* a = A + 3 * lane;
        if (!PN) {
            for (int c = 0; c < 3; c++) a[c] = a[c] + hdr[c] - tmp1 * xb[c];
        } else {
            double rv = xb[0] * vb[0] + xb[1] * vb[1] + xb[2] * vb[2];
            double v2 = vb[0] * vb[0] + vb[1] * vb[1] + vb[2] * vb[2];
            double f = tmp1 / light2;
            double radial = 4.0 * m[0] * s_1 - v2;
            for (int c = 0; c < 3; c++) a[c] = a[c] + hdr[c] - tmp1 * xb[c] + f * (radial * xb[c] + 4.0 * rv * vb[c]);
        }
        if (d.obl)
            for (int c = 0; c < 3; c++) a[c] = a[c] + (aob[3 * lane + c] - hdr[3 + c]);
        if (NG) {
            double pb[5];
            for (int q = 0; q < 5; q++) pb[q] = ngf[q * cap + d.base + 1 + lane];
            add_nongrav(d.ngflag, pb, m[lane + 1], xb[0], xb[1], xb[2], vb[0], vb[1], vb[2], a[0], a[1], a[2], fault);
        }
        for (int c = 0; c < 3; c++) T[kHdr + kF * lane + c] = xb[c];
        T[kHdr + kF * lane + 3] = m[lane + 1];
    }
    __syncthreads();
}

// NG: non-gravitational model on (d.ngflag != 0), which also turns the
// velocity predictor on (mercury6_2.for:4276-4292).
template <bool PN, bool NG>
__global__ void __launch_bounds__(64) ra_planets(RaArgs args) {
    const Sys d = args.sys[0];
    const int np = d.nbig - 1;
    const int lane = threadIdx.x;
    const long long cap = args.st.cap;
    __shared__ double m[kMaxBig], X[3 * kMaxBig], V[3 * kMaxBig], A[3 * kMaxBig], aob[3 * kMaxBig], hdr[8];
    __shared__ double X1[3 * kMaxBig], V1[3 * kMaxBig], A1[3 * kMaxBig];
    __shared__ double ps3[kPairMax], pdx[3 * kPairMax], prc[kMaxBig];
    __shared__ unsigned char ppi[kPairMax], ppj[kPairMax];
    for (int q = lane; q < d.nbig; q += 64) m[q] = args.st.m[d.base + q];
    if (lane == 0 && np <= kPairPlanets)
        for (int i = 0, q = 0; i < np; i++)
            for (int j = i + 1; j < np; j++, q++) { ppi[q] = (unsigned char)i; ppj[q] = (unsigned char)j; }
    const RaPairs pairs{ps3, pdx, prc, ppi, ppj};
    const RaPairs* P = np <= kPairPlanets ? &pairs : nullptr;
    double b[3][7], g[3][7];
    const int gb = d.base + 1 + lane;
    const bool iter = args.niter > 0;
    if (lane < np) {
        for (int c = 0; c < 3; c++) {
            double x1 = args.st.x[c * cap + gb], v1 = args.st.v[c * cap + gb];
            if (args.pend) {
                double e[7];
                for (int q = 0; q < 7; q++) {
                    b[c][q] = args.ra.b[(c * 7 + q) * cap + gb];
                    e[q] = args.ra.e[(c * 7 + q) * cap + gb];
                }
                ra_finish_comp(args.ptdid, args.pq, args.preset, args.reset, b[c], e,
                               args.ra.a1[c * cap + gb], x1, v1);
                for (int q = 0; q < 7; q++) args.ra.e[(c * 7 + q) * cap + gb] = e[q];
                args.st.x[c * cap + gb] = x1;
                args.st.v[c * cap + gb] = v1;
            } else if (iter) {
                for (int q = 0; q < 7; q++) {
                    b[c][q] = args.reset ? 0.0 : args.ra.b[(c * 7 + q) * cap + gb];
                    if (args.reset) args.ra.e[(c * 7 + q) * cap + gb] = 0.0;
                }
            }
            X1[3 * lane + c] = x1;
            V1[3 * lane + c] = v1;
            X[3 * lane + c] = x1;
            V[3 * lane + c] = v1;
        }
    }
    __syncthreads();
    ra_planet_forces<PN, NG>(d, np, m, X, V, A, hdr, aob, args.light2, args.tab, args.st.fault, P, args.st.ngf, cap);
    if (lane < np) {
        for (int c = 0; c < 3; c++) {
            A1[3 * lane + c] = A[3 * lane + c];
            if (iter) ra_g_from_b(b[c], g[c]);
        }
    }
    int entry = 1;
    for (int n = 1; n <= args.niter; n++) {
#pragma unroll 1
        for (int j = 2; j <= 8; j++, entry++) {
            const double hj = args.k.h[j - 1];
            if (lane < np) {
                double s[9];
                ra_sx(args.t, hj, s);
                for (int c = 0; c < 3; c++)
                    X[3 * lane + c] = ra_xpred(s, b[c], A1[3 * lane + c], V1[3 * lane + c], X1[3 * lane + c]);
                if (PN || NG) {
                    ra_sv(args.t, hj, s);
                    for (int c = 0; c < 3; c++) V[3 * lane + c] = ra_vpred(s, b[c], A1[3 * lane + c], V1[3 * lane + c]);
                }
            }
            __syncthreads();
            ra_planet_forces<PN, NG>(d, np, m, X, V, A, hdr, aob, args.light2, args.tab + (long long)entry * d.ts,
                                     args.st.fault, P, args.st.ngf, cap);
            if (lane < np)
                for (int c = 0; c < 3; c++) ra_update(j, A[3 * lane + c], A1[3 * lane + c], g[c], b[c]);
        }
    }
    double bm = 0.0;
    if (lane < np) {
        for (int c = 0; c < 3; c++) {
            args.ra.a1[c * cap + gb] = A1[3 * lane + c];
            if (iter) {
                for (int q = 0; q < 7; q++) args.ra.b[(c * 7 + q) * cap + gb] = b[c][q];
                bm = fmax(bm, fabs(b[c][6]));
            }
        }
    }
    if (iter) merge_bmax(args.bmax, bm, args.st.fault);
}

// Uncapped registers (151-153 VGPRs, 3 waves per SIMD): caps for 4 or 5 waves
// spill b and g and measured 1.2x and 2.4x slower.
// NG: each particle's non-gravitational terms added after its gravity
// (particles are massless: upload refuses Small bodies with mass).
template <bool PN, bool NG>
__global__ void __launch_bounds__(128) ra_particles(RaArgs args) {
    const Sys d = args.sys[0];
    const int jb = d.nbig + blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = jb < d.n;
    const long long cap = args.st.cap;
    const int g_ = d.base + jb;
    const int np = d.nbig - 1;
    double x1[3] = {0, 0, 0}, v1[3] = {0, 0, 0}, a1[3] = {0, 0, 0};
    double b[3][7], g[3][7];
    double bm = 0.0;
    const bool iter = args.niter > 0;
    if (live) {
        for (int c = 0; c < 3; c++) {
            x1[c] = args.st.x[c * cap + g_];
            v1[c] = args.st.v[c * cap + g_];
            if (args.pend) {
                double e[7];
                for (int q = 0; q < 7; q++) {
                    b[c][q] = args.ra.b[(c * 7 + q) * cap + g_];
                    e[q] = args.ra.e[(c * 7 + q) * cap + g_];
                }
                ra_finish_comp(args.ptdid, args.pq, args.preset, args.reset, b[c], e,
                               args.ra.a1[c * cap + g_], x1[c], v1[c]);
                for (int q = 0; q < 7; q++) args.ra.e[(c * 7 + q) * cap + g_] = e[q];
                args.st.x[c * cap + g_] = x1[c];
                args.st.v[c * cap + g_] = v1[c];
            } else if (iter) {
                for (int q = 0; q < 7; q++) {
                    b[c][q] = args.reset ? 0.0 : args.ra.b[(c * 7 + q) * cap + g_];
                    if (args.reset) args.ra.e[(c * 7 + q) * cap + g_] = 0.0;
                }
            }
        }
        double pb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (NG) load_ngf(args.st, g_, pb);
        grav_small<PN, RangeOps>(d, const_tab(args.tab), np, args.light2, x1[0], x1[1], x1[2], v1[0], v1[1],
                                 v1[2], a1[0], a1[1], a1[2], args.st.fault);
        if (NG)
            add_nongrav(d.ngflag, pb, 0.0, x1[0], x1[1], x1[2], v1[0], v1[1], v1[2], a1[0], a1[1], a1[2],
                        args.st.fault);
        if (iter)
            for (int c = 0; c < 3; c++) ra_g_from_b(b[c], g[c]);
        int entry = 1;
        for (int n = 1; n <= args.niter; n++) {
#pragma unroll 1
            for (int j = 2; j <= 8; j++, entry++) {
                const double hj = args.k.h[j - 1];
                double s[9], x[3], v[3] = {0, 0, 0};
                ra_sx(args.t, hj, s);
                for (int c = 0; c < 3; c++) x[c] = ra_xpred(s, b[c], a1[c], v1[c], x1[c]);
                if (PN || NG) {
                    ra_sv(args.t, hj, s);
                    for (int c = 0; c < 3; c++) v[c] = ra_vpred(s, b[c], a1[c], v1[c]);
                }
                double a[3];
                grav_small<PN, RangeOps>(d, const_tab(args.tab + (long long)entry * d.ts), np, args.light2, x[0],
                                         x[1], x[2], v[0], v[1], v[2], a[0], a[1], a[2], args.st.fault);
                if (NG)
                    add_nongrav(d.ngflag, pb, 0.0, x[0], x[1], x[2], v[0], v[1], v[2], a[0], a[1], a[2],
                                args.st.fault);
                for (int c = 0; c < 3; c++) ra_update(j, a[c], a1[c], g[c], b[c]);
            }
        }
        for (int c = 0; c < 3; c++) {
            args.ra.a1[c * cap + g_] = a1[c];
            if (iter) {
                for (int q = 0; q < 7; q++) args.ra.b[(c * 7 + q) * cap + g_] = b[c][q];
                bm = fmax(bm, fabs(b[c][6]));
            }
        }
    }
    if (iter) merge_bmax(args.bmax, bm, args.st.fault);
}

// End of the sequence for every body j >= 1 (mercury6_2.for:4426-4489): the
// step-start state is kept in ox, ov (impact redo, event screen).
struct RaFinishArgs {
    const Sys* sys;
    State st;
    RaState ra;
    RaConst k;
    double tdid, q;
    int reset;
};

__global__ void __launch_bounds__(128) ra_finish(RaFinishArgs args) {
    const Sys d = args.sys[0];
    const int jb = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (jb >= d.n) return;
    const long long cap = args.st.cap;
    const int g = d.base + jb;
    const double tdid = args.tdid, temp = tdid * tdid;
    const double q = args.q, q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q3 * q4;
    constexpr RaConst K = make_ra_constants();
    const double* xc = K.xc;
    const double* vc = K.vc;
    for (int c = 0; c < 3; c++) {
        double b[7], e[7];
        for (int i = 0; i < 7; i++) {
            b[i] = args.ra.b[(c * 7 + i) * cap + g];
            e[i] = args.ra.e[(c * 7 + i) * cap + g];
        }
        double a1 = args.ra.a1[c * cap + g];
        double x1 = args.st.x[c * cap + g], v1 = args.st.v[c * cap + g];
        args.st.ox[c * cap + g] = x1;
        args.st.ov[c * cap + g] = v1;
        x1 = (xc[7] * b[6] + xc[6] * b[5] + xc[5] * b[4] + xc[4] * b[3] + xc[3] * b[2] + xc[2] * b[1] + xc[1] * b[0]
              + xc[0] * a1) * temp + v1 * tdid + x1;
        v1 = (vc[6] * b[6] + vc[5] * b[5] + vc[4] * b[4] + vc[3] * b[3] + vc[2] * b[2] + vc[1] * b[1] + vc[0] * b[0]
              + a1) * tdid + v1;
        args.st.x[c * cap + g] = x1;
        args.st.v[c * cap + g] = v1;
        double s[7];
        for (int i = 0; i < 7; i++) s[i] = args.reset ? 0.0 : b[i] - e[i];
        e[0] = q * (b[6] * 7.0 + b[5] * 6.0 + b[4] * 5.0 + b[3] * 4.0 + b[2] * 3.0 + b[1] * 2.0 + b[0]);
        e[1] = q2 * (b[6] * 21.0 + b[5] * 15.0 + b[4] * 10.0 + b[3] * 6.0 + b[2] * 3.0 + b[1]);
        e[2] = q3 * (b[6] * 35.0 + b[5] * 20.0 + b[4] * 10.0 + b[3] * 4.0 + b[2]);
        e[3] = q4 * (b[6] * 35.0 + b[5] * 15.0 + b[4] * 5.0 + b[3]);
        e[4] = q5 * (b[6] * 21.0 + b[5] * 6.0 + b[4]);
        e[5] = q6 * (b[6] * 7.0 + b[5]);
        e[6] = q7 * b[6];
        for (int i = 0; i < 7; i++) {
            args.ra.e[(c * 7 + i) * cap + g] = e[i];
            args.ra.b[(c * 7 + i) * cap + g] = e[i] + s[i];
        }
    }
}

// Planet screening block (mce_box, padded by 1.2 rce) for the event screen of
// a non-MVS step: start (ox, ov) and end (x, v) states of every Big body.
__global__ void __launch_bounds__(64) planet_screen_block(const Sys* sys, State st, double* T, double h) {
    const Sys d = sys[0];
    const int np = d.nbig - 1;
    const int p = threadIdx.x;
    const long long cap = st.cap;
    if (p >= np) return;
    const int g = d.base + 1 + p;
    double* B = T + kHdr + kF * np + kScr * p;
    double lo, hi;
    box_axis(h, st.ox[g], st.ov[g], st.x[g], st.v[g], lo, hi);
    double pad = st.rce[g] * 1.2;
    B[SX0] = lo - pad; B[SX1] = hi + pad;
    box_axis(h, st.ox[cap + g], st.ov[cap + g], st.x[cap + g], st.v[cap + g], lo, hi);
    B[SY0] = lo - pad; B[SY1] = hi + pad;
    for (int c = 0; c < 3; c++) {
        B[SP + c] = st.ox[c * cap + g];
        B[SV0 + c] = st.ov[c * cap + g];
        B[SV1 + c] = st.v[c * cap + g];
        T[kHdr + kF * p + c] = st.x[c * cap + g];
    }
    B[SRCE] = st.rce[g];
    B[SRPH] = st.rphys[g];
    T[kHdr + kF * p + 3] = st.m[g];
}

}  // namespace mhip
