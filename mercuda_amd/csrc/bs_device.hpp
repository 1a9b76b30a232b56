// Bulirsch-Stoer integration of a whole system for gfx950, FP64: algorithm 2
// (mdt_bs1, mercury6_2.for:3227-3426) and algorithm 3 (mdt_bs2,
// mercury6_2.for:3428-3617), each with mfo_all as the force (mfo_grav with the
// optional central 1PN term, plus oblateness and the non-gravitational model;
// mercury6_2.for:4506-4707, 5106-5357).
//
// Massless particles do not act on the Big bodies, so one attempt runs as
//   bsw_big        one block, a thread per (level, planet): every
//                  modified-midpoint level advances in lockstep, one force
//                  evaluation per round, and each evaluation tabulates what a
//                  particle's force at that point needs (indirect term,
//                  oblateness reaction, planet positions and masses, the RA15
//                  table layout); then the planets' extrapolation, level
//                  errors and accepted states;
//   bsw_particles  a thread per (level, particle), the force read from those
//                  tables; the same extrapolation per particle;
//   bsw_accept     the level the host chose (first n > 3 with the largest
//                  error within tol^2), the step-start states kept in ox, ov.
// The levels are independent until the extrapolation, which replays the
// Fortran's order per (body, component), so the bits are the sequential
// loop's.  mdt_bs1 passes each point's own velocities to the force, mdt_bs2
// the step-start velocities v0 (its force is meant to be conservative).
#pragma once
#include "radau_device.hpp"

namespace mhip {

constexpr int kBsLanes = 16;      // planets one level of bsw_big holds (up to 16 planets)
constexpr int kBsWide = 64;       // the same for more planets (up to kMaxBig - 1)
constexpr int kBsMaxLevels = 12;

template <int ALG> struct BsShape;
template <> struct BsShape<2> { static constexpr int L = 8; static constexpr int R = 16; static constexpr int P = 72; };
template <> struct BsShape<3> { static constexpr int L = 12; static constexpr int R = 12; static constexpr int P = 78; };

// force evaluations of level n, and the table slot of its r-th (slot 0: the
// step start; then the evaluations in the sequential order)
template <int ALG> __host__ __device__ constexpr int bsw_evals(int n) { return ALG == 2 ? 2 * n : n; }
template <int ALG> __host__ __device__ constexpr int bsw_slot(int n, int r) {
    return 1 + (ALG == 2 ? n * (n - 1) : n * (n - 1) / 2) + r - 1;
}

struct BswArgs {
    const Sys* sys;
    State st;
    double* tab;                // 1 + P tables of stride Sys::ts
    double* cand;               // accepted states per level n > 3: [L - 3][6][cap]
    double* rsave;              // the first pass's level results for the second: [8][6][cap]
    unsigned long long* err;    // max error per level [12] (bits of non-negative doubles)
    double h0;
    double light2;
};

// One level of one body: h2 of the extrapolation, then the running sums.
// ALG 2: x, v and xe, ve (xend, vend); ALG 3: b, c and xe (xend).
template <int ALG>
struct BswLevel {
    double h, hby2, h2, h2by2, hx2;
    double x[3], v[3], xe[3], ve[3];
};

template <int ALG>
__device__ __forceinline__ double bsw_h2(double h0, int n) {
    if (ALG == 2) return .25 / (double)(n * n);
    double h = h0 / (double)n;
    return h * h;
}

template <int ALG>
__device__ __forceinline__ void bsw_start(BswLevel<ALG>& L, int n, double h0, const double* x0, const double* v0,
                                          const double* a0) {
    if (ALG == 2) {
        L.h = h0 / (2.0 * (double)n);
        L.hx2 = L.h * 2.0;
        for (int q = 0; q < 3; q++) {
            L.x[q] = x0[q] + L.h * v0[q];
            L.v[q] = v0[q] + L.h * a0[q];
        }
    } else {
        L.h = h0 / (double)n;
        L.hby2 = .5 * L.h;
        L.h2 = L.h * L.h;
        L.h2by2 = .5 * L.h2;
        for (int q = 0; q < 3; q++) {
            L.x[q] = .5 * a0[q];       // b
            L.v[q] = 0.0;              // c
            L.xe[q] = L.h2by2 * a0[q] + L.h * v0[q] + x0[q];
        }
    }
}

// Where evaluation r of level n is taken: position and the velocity the force sees.
template <int ALG>
__device__ __forceinline__ void bsw_point(const BswLevel<ALG>& L, int r, const double* v0, const double*& xp,
                                          const double*& vp) {
    if (ALG == 2) {
        if (r & 1) { xp = L.x; vp = L.v; }
        else { xp = L.xe; vp = L.ve; }
    } else {
        xp = L.xe; vp = v0;
    }
}

// After the force a of evaluation r < evals(n): advance to the next point.
template <int ALG>
__device__ __forceinline__ void bsw_advance(BswLevel<ALG>& L, int r, const double* a, const double* x0,
                                            const double* v0, const double* a0) {
    if (ALG == 2) {
        if (r == 1) {
            for (int q = 0; q < 3; q++) {
                L.xe[q] = x0[q] + L.hx2 * L.v[q];
                L.ve[q] = v0[q] + L.hx2 * a[q];
            }
        } else if (r & 1) {
            for (int q = 0; q < 3; q++) {
                L.xe[q] = L.xe[q] + L.hx2 * L.v[q];
                L.ve[q] = L.ve[q] + L.hx2 * a[q];
            }
        } else {
            for (int q = 0; q < 3; q++) {
                L.x[q] = L.x[q] + L.hx2 * L.ve[q];
                L.v[q] = L.v[q] + L.hx2 * a[q];
            }
        }
    } else {
        const int j = r + 1;
        double tmp0 = L.h * (double)j;
        for (int q = 0; q < 3; q++) {
            L.x[q] = L.x[q] + a[q];
            L.v[q] = L.v[q] + L.x[q];
            L.xe[q] = L.h2 * L.v[q] + L.h2by2 * a0[q] + tmp0 * v0[q] + x0[q];
        }
    }
}

// The level's result from the force a at its last evaluation.
template <int ALG>
__device__ __forceinline__ void bsw_result(const BswLevel<ALG>& L, const double* a, const double* v0, double* R) {
    if (ALG == 2) {
        for (int q = 0; q < 3; q++) {
            R[q] = .5 * (L.xe[q] + L.x[q] + L.h * L.ve[q]);
            R[3 + q] = .5 * (L.ve[q] + L.v[q] + L.h * a[q]);
        }
    } else {
        for (int q = 0; q < 3; q++) {
            R[q] = L.xe[q];
            R[3 + q] = L.h * L.x[q] + L.hby2 * a[q] + v0[q];
        }
    }
}

// The extrapolation weights of mdt_bs1 / mdt_bs2 for the step h0: for level n
// and j = n-1 .. 1, tmp0 = 1/(h2(j) - h2(n)) and W = (tmp0 * h2(j+1), tmp0 *
// h2(n)).  They depend on h0 alone, so a block forms them once (threads t of
// nt; a barrier must follow) instead of every thread per component.
template <int ALG> __host__ __device__ constexpr int bsw_pairs() { return BsShape<ALG>::L * (BsShape<ALG>::L - 1) / 2; }
__host__ __device__ constexpr int bsw_ex(int n, int j) { return (n - 1) * (n - 2) / 2 + j - 1; }
template <int ALG>
__device__ inline void bsw_weights(double h0, double (*W)[2], int t, int nt) {
    for (int w = t; w < bsw_pairs<ALG>(); w += nt) {
        int n = 2;
        while (n * (n - 1) / 2 <= w) n++;
        const int j = w - (n - 1) * (n - 2) / 2 + 1;
        const double h2n = bsw_h2<ALG>(h0, n), h2j = bsw_h2<ALG>(h0, j), h2j1 = bsw_h2<ALG>(h0, j + 1);
        const double tmp0 = 1.0 / (h2j - h2n);
        W[w][0] = tmp0 * h2j1;
        W[w][1] = tmp0 * h2n;
    }
}

// Extrapolation of one component of one body over the levels 1..NMAX in the
// Fortran's order, every level's result already in D (row n-1): per level
// n > 3 and n >= nlo, d(1) in E[n-4] and the accepted value in cand (a
// second pass redoes the first levels' updates, whose outputs it already has).
template <int NMAX>
__device__ inline void bsw_extrapolate(const double (*W)[2], double* D, int ds, double* E, double* cand,
                                       long long cs, int nlo) {
    for (int n = 1; n <= NMAX; n++) {
        for (int j = n - 1; j >= 1; j--) {
            const int j1 = j + 1;
            const double tmp1 = W[bsw_ex(n, j)][0];
            const double tmp2 = W[bsw_ex(n, j)][1];
            D[(j - 1) * 6 * ds] = tmp1 * D[(j1 - 1) * 6 * ds] - tmp2 * D[(j - 1) * 6 * ds];
        }
        if (n > 3 && n >= nlo) {
            double d1 = D[0];
            E[n - 4] = d1;
            double s = d1;
            for (int j = 2; j <= n; j++) s = s + D[(j - 1) * 6 * ds];
            cand[(long long)(n - 4) * 6 * cs] = s;
        }
    }
}

// Level errors of one body (mdt_bs1 mercury6_2.for:3348-3362 / mdt_bs2), merged by max.
__device__ __forceinline__ double bsw_level_err(const double* d1, double xs, double vs, int* fault) {
    bool fin = true;
    for (int q = 0; q < 6; q++) fin = fin && isfinite(d1[q]);
    double tmp1 = fmax(fmax(d1[0] * d1[0], d1[1] * d1[1]), d1[2] * d1[2]);
    double tmp2 = fmax(fmax(d1[3] * d1[3], d1[4] * d1[4]), d1[5] * d1[5]);
    if (!fin) atomicOr(fault, FAULT_FORCE);
    return fin ? fmax(tmp1 * xs, tmp2 * vs) : 0.0;
}

__device__ __forceinline__ void bsw_scales(const double* x0, const double* v0, double& xs, double& vs) {
    double t1 = x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2];
    double t2 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
    xs = 1.0 / fmax(t1, 1e-300);
    vs = 1.0 / fmax(t2, 1e-300);
}

// Big-body accelerations of mfo_all (mfo_grav mercury6_2.for:4616-4690, 1PN,
// oblateness) for one level's planets at X, V ([3][lanes] each), the
// arithmetic of ra_planet_forces: thread (level, k) computes planet k+1; lane 0
// of the level forms the indirect term and the oblateness reaction.  When act,
// the level's table T for this point is written too.  Every thread of the
// block calls it (block barriers inside).
// NG: the non-gravitational terms added last (ngf: [5][cap] parameters).
template <bool PN, bool NG>
__device__ void bsw_planet_forces(const Sys& d, int np, const double* m, const double* X, const double* V, double* A,
                                  double* aob, double* hdr, double light2, double* T, int* fault, int k, bool act,
                                  const double* ngf, long long cap) {
    if (act && k < np) {
        const double* xb = X + 3 * k;
        double a[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < k; i++) {
            double dx = xb[0] - X[3 * i], dy = xb[1] - X[3 * i + 1], dz = xb[2] - X[3 * i + 2];
            double s2 = dx * dx + dy * dy + dz * dz;
            if (!(s2 > 0.0)) atomicOr(fault, FAULT_FORCE);
            double s_1 = 1.0 / sqrt(s2);
            double s_3 = s_1 * s_1 * s_1;
            double tmp1 = s_3 * m[i + 1];
            a[0] = a[0] - tmp1 * dx; a[1] = a[1] - tmp1 * dy; a[2] = a[2] - tmp1 * dz;
        }
        for (int j = k + 1; j < np; j++) {
            double dx = X[3 * j] - xb[0], dy = X[3 * j + 1] - xb[1], dz = X[3 * j + 2] - xb[2];
            double s2 = dx * dx + dy * dy + dz * dz;
            if (!(s2 > 0.0)) atomicOr(fault, FAULT_FORCE);
            double s_1 = 1.0 / sqrt(s2);
            double s_3 = s_1 * s_1 * s_1;
            double tmp2 = s_3 * m[j + 1];
            a[0] = a[0] + tmp2 * dx; a[1] = a[1] + tmp2 * dy; a[2] = a[2] + tmp2 * dz;
        }
        for (int c = 0; c < 3; c++) A[3 * k + c] = a[c];
        if (d.obl) obl_body(d.mu, d.j2, d.j4, d.j6, xb[0], xb[1], xb[2], aob[3 * k], aob[3 * k + 1], aob[3 * k + 2]);
    }
    __syncthreads();
    if (act && k == 0) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int i = 0; i < np; i++) {
            const double* xi = X + 3 * i;
            double s2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
            if (!(s2 > 0.0)) atomicOr(fault, FAULT_FORCE);
            double s_1 = 1.0 / sqrt(s2);
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
        for (int q = 0; q < 6; q++) T[q] = hdr[q];
        T[6] = 0.0; T[7] = 0.0;
    }
    __syncthreads();
    if (act && k < np) {
        const double* xb = X + 3 * k;
        const double* vb = V + 3 * k;
        double s2 = xb[0] * xb[0] + xb[1] * xb[1] + xb[2] * xb[2];
        double s_1 = 1.0 / sqrt(s2);
        double r3 = s_1 * s_1 * s_1;
        double tmp1 = m[0] * r3;
        double* a = A + 3 * k;
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
            for (int c = 0; c < 3; c++) a[c] = a[c] + (aob[3 * k + c] - hdr[3 + c]);
        if (NG) {
            double pb[5];
            for (int q = 0; q < 5; q++) pb[q] = ngf[q * cap + d.base + 1 + k];
            add_nongrav(d.ngflag, pb, m[k + 1], xb[0], xb[1], xb[2], vb[0], vb[1], vb[2], a[0], a[1], a[2], fault);
        }
        for (int c = 0; c < 3; c++) T[kHdr + kF * k + c] = xb[c];
        T[kHdr + kF * k + 3] = m[k + 1];
    }
    __syncthreads();
}

// One attempt over the levels LO..HI, planets: one block of (HI-LO+1) x NB
// threads, thread (level n, planet k+1); NB = kBsLanes lanes per level for up
// to 16 planets, kBsWide for more.  An attempt runs levels
// 1..HI first; the host runs the rest (LO > 1) only when none of those levels
// converged.  A first pass (HI < L) keeps its levels' results in rsave; the
// second reads them back, so that its extrapolation is that of all L levels.
template <int ALG, bool PN, bool NG, int LO, int HI, int NB>
__global__ void __launch_bounds__((HI - LO + 1) * NB) bsw_big(BswArgs A) {
    constexpr int L = BsShape<ALG>::L;
    constexpr int NL = HI - LO + 1;
    constexpr int NT = NL * NB;
    constexpr int R = bsw_evals<ALG>(HI);
    constexpr int ELO = LO > 4 ? LO : 4;
    static_assert(LO == 1 || HI == L, "a first pass or the completing pass");
    static_assert(LO - 1 <= 8, "rsave holds 8 levels");
    const Sys d = A.sys[0];
    const int np = d.nbig - 1;
    const int t = threadIdx.x;
    const int n = LO + t / NB;
    const int k = t % NB;
    const bool mem = k < np;
    const long long cap = A.st.cap;
    __shared__ double m[NB + 1];
    __shared__ double X[NL][3 * NB], V[NL][3 * NB], F[NL][3 * NB], OB[NL][3 * NB],
        HD[NL][8];
    __shared__ double D[6 * HI * NB];
    __shared__ double E[NB][6][L - 3];
    __shared__ double W[bsw_pairs<ALG>()][2];
    if (t <= np) m[t] = A.st.m[d.base + t];
    bsw_weights<ALG>(A.h0, W, t, NT);
    if constexpr (LO > 1) {
        for (int w = t; w < (LO - 1) * 6 * NB; w += NT) {
            const int row = w / (6 * NB), q = w / NB % 6, km = w % NB;
            if (km < np) D[(6 * row + q) * NB + km] = A.rsave[(long long)(6 * row + q) * cap + d.base + 1 + km];
        }
    }
    double x0[3] = {0, 0, 0}, v0[3] = {0, 0, 0}, a0[3] = {0, 0, 0};
    if (mem)
        for (int c = 0; c < 3; c++) {
            x0[c] = A.st.x[c * cap + d.base + 1 + k];
            v0[c] = A.st.v[c * cap + d.base + 1 + k];
            if (n == LO) { X[0][3 * k + c] = x0[c]; V[0][3 * k + c] = v0[c]; }
        }
    __syncthreads();
    // the step start: one evaluation (the first level's lanes), shared by every level
    bsw_planet_forces<PN, NG>(d, np, m, X[0], V[0], F[0], OB[0], HD[0], A.light2, A.tab + d.tab, A.st.fault, k, n == LO,
                              A.st.ngf, cap);
    if (mem)
        for (int c = 0; c < 3; c++) a0[c] = F[0][3 * k + c];
    __syncthreads();
    BswLevel<ALG> Lv;
    bsw_start<ALG>(Lv, n, A.h0, x0, v0, a0);
    double a[3] = {0, 0, 0};
    for (int r = 1; r <= R; r++) {
        const bool act = r <= bsw_evals<ALG>(n);
        if (act && mem) {
            const double *xp, *vp;
            bsw_point<ALG>(Lv, r, v0, xp, vp);
            for (int c = 0; c < 3; c++) { X[n - LO][3 * k + c] = xp[c]; V[n - LO][3 * k + c] = vp[c]; }
        }
        __syncthreads();
        bsw_planet_forces<PN, NG>(d, np, m, X[n - LO], V[n - LO], F[n - LO], OB[n - LO], HD[n - LO], A.light2,
                                  A.tab + d.tab + (long long)bsw_slot<ALG>(n, act ? r : 1) * d.ts, A.st.fault, k, act,
                                  A.st.ngf, cap);
        if (act && mem) {
            for (int c = 0; c < 3; c++) a[c] = F[n - LO][3 * k + c];
            if (r < bsw_evals<ALG>(n)) bsw_advance<ALG>(Lv, r, a, x0, v0, a0);
        }
    }
    if (mem) {
        double Rr[6];
        bsw_result<ALG>(Lv, a, v0, Rr);
        for (int q = 0; q < 6; q++) D[(6 * (n - 1) + q) * NB + k] = Rr[q];
        if constexpr (HI < L)
            for (int q = 0; q < 6; q++) A.rsave[(long long)(6 * (n - 1) + q) * cap + d.base + 1 + k] = Rr[q];
    }
    // extrapolation: item (planet, component), in passes separated by barriers
    for (int p = 0; p < (6 * NB + NT - 1) / NT; p++) {
        __syncthreads();
        const int w = p * NT + t;
        const int km = w / 6, q = w % 6;
        if (w < 6 * NB && km < np) {
            double e[L - 3];
            bsw_extrapolate<HI>(W, D + q * NB + km, NB, e, A.cand + (long long)q * cap + d.base + 1 + km,
                                cap, LO);
            for (int l = ELO - 4; l <= HI - 4; l++) E[km][q][l] = e[l];
        }
    }
    __syncthreads();
    constexpr int NE = HI - ELO + 1;
    for (int w = t; w < NE * NB; w += NT) {
        const int km = w / NE, l = ELO - 4 + w % NE;
        if (km < np) {
            double xm[3], vm[3], xs, vs, d1[6];
            for (int c = 0; c < 3; c++) {
                xm[c] = A.st.x[c * cap + d.base + 1 + km];
                vm[c] = A.st.v[c * cap + d.base + 1 + km];
            }
            bsw_scales(xm, vm, xs, vs);
            for (int q = 0; q < 6; q++) d1[q] = E[km][q][l];
            double err = bsw_level_err(d1, xs, vs, A.st.fault);
            atomicMax(A.err + 3 + l, (unsigned long long)__double_as_longlong(err));
        }
    }
}

// One attempt, particles (after bsw_big).  Level n makes evals(n) force
// evaluations (2n for ALG 2, n for ALG 3), so thread (pair p, particle slot)
// runs the levels p+1 and L-p one after the other: every thread then makes the
// same number of evaluations (2(L+1) or L+1) and no lane idles while a longer
// level of its wave runs.  The force is the RA15 particle force on the tables
// bsw_big wrote at the same points; each level's arithmetic and its d row are
// those of a thread of its own.
// bsw_particles<ALG, PN, LO, HI> runs the levels LO..HI with bsw_big's passes
// and rsave convention: a first pass in pairs (LO+p, HI-p), the completing
// pass one level per thread (it holds fewer levels, and its blocks need more
// threads per slot of LDS to stay resident).
constexpr int kBswSlots = 32;      // particles per block of bsw_particles
template <int LO, int HI> __host__ __device__ constexpr bool bsw_paired() { return LO == 1; }
template <int LO, int HI> __host__ __device__ constexpr int bsw_threads() {
    return (bsw_paired<LO, HI>() ? (HI - LO + 1) / 2 : HI - LO + 1) * kBswSlots;
}

template <int ALG, bool PN, bool NG, int LO, int HI>
__global__ void __launch_bounds__((bsw_threads<LO, HI>())) bsw_particles(BswArgs A) {
    constexpr int L = BsShape<ALG>::L;
    constexpr int NS = kBswSlots;
    constexpr int NT = bsw_threads<LO, HI>();
    constexpr int ELO = LO > 4 ? LO : 4;
    constexpr bool PAIR = bsw_paired<LO, HI>();
    static_assert(!PAIR || (HI - LO + 1) % 2 == 0, "levels pair up");
    static_assert(LO == 1 || HI == L, "a first pass or the completing pass");
    const Sys d = A.sys[0];
    const int np = d.nbig - 1;
    const int t = threadIdx.x;
    const int pair = t / NS;
    const int sl = t % NS;
    const int npart = d.n - d.nbig;
    const int pi_ = blockIdx.x * NS + sl;
    const bool live = pi_ < npart;
    const long long cap = A.st.cap;
    const int g = d.base + d.nbig + pi_;
    __shared__ double D[6 * HI * NS];
    __shared__ double E[NS][6][L - 3];
    __shared__ double W[bsw_pairs<ALG>()][2];
    __shared__ unsigned long long EM[L - 3];
    if (t < L - 3) EM[t] = 0ull;
    bsw_weights<ALG>(A.h0, W, t, NT);
    if constexpr (LO > 1) {
        for (int w = t; w < (LO - 1) * 6 * NS; w += NT) {
            const int row = w / (6 * NS), q = w / NS % 6, s = w % NS;
            const int ps = blockIdx.x * NS + s;
            if (ps < npart) D[(6 * row + q) * NS + s] = A.rsave[(long long)(6 * row + q) * cap + d.base + d.nbig + ps];
        }
    }
    if (live) {
        double x0[3], v0[3], a0[3];
        for (int c = 0; c < 3; c++) { x0[c] = A.st.x[c * cap + g]; v0[c] = A.st.v[c * cap + g]; }
        double pb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (NG) load_ngf(A.st, g, pb);
        grav_small<PN, RangeOps>(d, A.tab + d.tab, np, A.light2, x0[0], x0[1], x0[2], v0[0], v0[1], v0[2], a0[0],
                                 a0[1], a0[2], A.st.fault);
        if (NG)
            add_nongrav(d.ngflag, pb, 0.0, x0[0], x0[1], x0[2], v0[0], v0[1], v0[2], a0[0], a0[1], a0[2], A.st.fault);
#pragma unroll 1
        for (int k = 0; k < (PAIR ? 2 : 1); k++) {
            const int n = k == 0 ? LO + pair : HI - pair;
            BswLevel<ALG> Lv;
            bsw_start<ALG>(Lv, n, A.h0, x0, v0, a0);
            double a[3];
            const int ne = bsw_evals<ALG>(n);
            for (int r = 1; r <= ne; r++) {
                const double *xp, *vp;
                bsw_point<ALG>(Lv, r, v0, xp, vp);
                grav_small<PN, RangeOps>(d, A.tab + d.tab + (long long)bsw_slot<ALG>(n, r) * d.ts, np, A.light2,
                                         xp[0], xp[1], xp[2], vp[0], vp[1], vp[2], a[0], a[1], a[2], A.st.fault);
                if (NG)
                    add_nongrav(d.ngflag, pb, 0.0, xp[0], xp[1], xp[2], vp[0], vp[1], vp[2], a[0], a[1], a[2],
                                A.st.fault);
                if (r < ne) bsw_advance<ALG>(Lv, r, a, x0, v0, a0);
            }
            double Rr[6];
            bsw_result<ALG>(Lv, a, v0, Rr);
            for (int q = 0; q < 6; q++) D[(6 * (n - 1) + q) * NS + sl] = Rr[q];
            if constexpr (HI < L)
                for (int q = 0; q < 6; q++) A.rsave[(long long)(6 * (n - 1) + q) * cap + g] = Rr[q];
        }
    }
    // extrapolation: work item (slot, component), in passes separated by
    // barriers (the weights are read from LDS in each pass, never kept in
    // registers across passes)
    for (int p = 0; p < (6 * NS + NT - 1) / NT; p++) {
        __syncthreads();
        const int w = p * NT + t;
        const int s = w / 6, q = w % 6;
        const int ps = blockIdx.x * NS + s;
        if (w < 6 * NS && ps < npart) {
            double e[L - 3];
            bsw_extrapolate<HI>(W, D + q * NS + s, NS, e, A.cand + (long long)q * cap + d.base + d.nbig + ps, cap, LO);
            for (int l = ELO - 4; l <= HI - 4; l++) E[s][q][l] = e[l];
        }
    }
    __syncthreads();
    // level errors: work item (slot, level), merged per level over the block
    constexpr int NE = HI - ELO + 1;
    for (int w = t; w < NE * NS; w += NT) {
        const int s = w / NE, l = ELO - 4 + w % NE;
        const int ps = blockIdx.x * NS + s;
        if (ps < npart) {
            const int gp = d.base + d.nbig + ps;
            double xm[3], vm[3], xs, vs, d1[6];
            for (int c = 0; c < 3; c++) { xm[c] = A.st.x[c * cap + gp]; vm[c] = A.st.v[c * cap + gp]; }
            bsw_scales(xm, vm, xs, vs);
            for (int q = 0; q < 6; q++) d1[q] = E[s][q][l];
            double err = bsw_level_err(d1, xs, vs, A.st.fault);
            atomicMax(&EM[l], (unsigned long long)__double_as_longlong(err));
        }
    }
    __syncthreads();
    if (t >= ELO - 4 && t <= HI - 4 && EM[t]) atomicMax(A.err + 3 + t, EM[t]);
}

// The accepted level's states for every body j >= 1, the step-start states
// kept in ox, ov (the driver's impact redo and the event screen).
__global__ void bsw_accept(BswArgs A, int level) {
    const Sys d = A.sys[0];
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j >= d.n) return;
    const long long cap = A.st.cap;
    const int g = d.base + j;
    for (int c = 0; c < 3; c++) {
        A.st.ox[c * cap + g] = A.st.x[c * cap + g];
        A.st.ov[c * cap + g] = A.st.v[c * cap + g];
        A.st.x[c * cap + g] = A.cand[(long long)((level - 4) * 6 + c) * cap + g];
        A.st.v[c * cap + g] = A.cand[(long long)((level - 4) * 6 + 3 + c) * cap + g];
    }
}

}  // namespace mhip
