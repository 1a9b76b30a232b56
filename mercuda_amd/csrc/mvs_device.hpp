// MVS step for massless particles on gfx950: data layout and kernels.
//
// One MVS step (mdt_mvs, mercury6_2.for:3987-4097) splits into
//   * a planet prologue: the Big bodies' kick, Jacobi drift and interaction
//     force.  Massless particles never act on Big bodies (mfo_mvs,
//     mercury6_2.for:5016-5054), so one wavefront per system advances the
//     planets S steps ahead and tabulates, per step, what every particle
//     needs: the indirect term a0tp, the oblateness reaction acen, the planet
//     positions after the drift with their masses, and the planets' screening
//     data (padded x-y boxes, start/end states).
//   * a particle pass: one lane per particle, SoA state in HBM, the planet
//     table read with wave-uniform (scalar) loads.  The fast kernel handles
//     the kepmd branch of the drift; a lane whose drift would take any other
//     branch parks (wave-ballot compaction, one atomic per wave) and the
//     general kernel finishes it from that step with the full drift chain.
// Both particle kernels run kick -> drift -> encounter screen (mce_stat) ->
// force (mfo_mvs small-body terms) -> kick -> central check (mce_cent) per
// step, with the Fortran's operation order, so results are bit-identical to
// the CPU path apart from libm calls.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "kepler.hpp"
#include "nongrav.hpp"

namespace mhip {

constexpr int kMaxBig = 64;            // Big bodies per system, central body included
constexpr int kMaxPpl = 8;             // particles per lane of the fast kernel
constexpr int kHdr = 8;                // table header: a0tp[3], acen[3], time_end, pad
constexpr int kF = 4;                  // force block per planet: X1[3], m
constexpr int kScr = 15;               // screen block per planet
// screen block layout per planet
enum { SX0 = 0, SX1, SY0, SY1, SP = 4, SV0 = 7, SV1 = 10, SRCE = 13, SRPH = 14 };

__host__ __device__ inline int table_stride(int np) { return (kHdr + (kF + kScr) * np + 7) & ~7; }

// A step table seen through the constant address space.  Uniform reads from
// it compile to scalar loads into SGPRs, also in loops that store to global
// memory (request and event records), where the compiler could not otherwise
// show the table unclobbered and would read it with vector loads into VGPRs.
// Only for tables that no thread writes while the kernel runs.
typedef const __attribute__((address_space(4))) double ConstTab;
__device__ __forceinline__ ConstTab* const_tab(const double* p) { return (ConstTab*)p; }
__host__ __device__ inline int start_stride(int np) { return (kHdr + kF * np + 7) & ~7; }

// Per-system launch parameters (device array, read with scalar loads).
// Static for a given upload, step size and launch length; the per-launch
// table base pointer and start time are kernel arguments.
struct Sys {
    int base;          // index of this system's central body in the SoA arrays
    int n;             // bodies in the system, central body included
    int nbig;          // Big bodies, central body included
    int ts;            // table stride (doubles) per step
    long long tab;     // offset (doubles) of this system's step tables in a launch's table block
    long long start;   // offset (doubles) of this system's start table (used when reset)
    double mu;         // G * central mass
    double h, hby2;    // step and half step
    double rcen;       // central-body radius (impact check)
    double j2, j4, j6; // central oblateness terms J_n R^n
    int obl;           // jcen != 0
    int ngflag;        // non-gravitational model: 1 cometary and Yarkovsky, 2 radiation, 3 both
};

// Body-state arrays: component-major SoA over all bodies of all systems.
struct State {
    double* x;     // [3][cap]
    double* v;     // [3][cap]
    double* a;     // [3][cap]  accelerations carried between steps
    const double* m;
    const double* rce;
    const double* rphys;
    // step-start copies for the host-driven event call (single-step mode only)
    double* ox;    // position before the step
    double* ov;    // velocity before the step
    double* v0k;   // velocity after the first kick
    double* v1;    // velocity after the drift
    int* fault;    // set on a non-finite acceleration
    long long cap;
    const double* ngf = nullptr;   // [5][cap] A1, A2, A3, beta, yar (non-gravitational model only)
    double* angf = nullptr;        // [3][cap] MVS: cometary accelerations carried between steps
};

// The five non-gravitational parameters of body g.
__device__ __forceinline__ void load_ngf(const State& st, int g, double P[5]) {
    for (int q = 0; q < 5; q++) P[q] = st.ngf[q * st.cap + g];
}

// Device event record: the reference's 128-byte MercuryEvent (mercury_cuda.h:7-11)
// followed by the step within the launch and the system index.
struct DevEvent {
    int i, j, kind, now;
    double distance, time;
    double xi[6], xj[6];
    int step, sys;
};

// Fault and overflow flags (bit set in State::fault / the sinks' flags word).
enum { FAULT_FORCE = 1, FAULT_REQ_OVERFLOW = 2, FAULT_EVENT_OVERFLOW = 4 };

struct EventSink {
    DevEvent* buf;
    int* count;
    int cap;
    int* flags;
};

// A member of the slow set: a particle the general kernel advances with the
// whole drift chain, from `step` of launch `chunk` of the current call and from
// step 0 of every later launch.  Members kept from the previous call carry
// chunk -1 (every step of the call).
struct Park {
    int j, step, sys, chunk;
};

// The slow set persists across calls, so that a particle parked by the fast
// kernel in one call starts the next one in the general kernel instead of
// parking again after the fast kernel's first launch.  Two lists of `cap`
// entries alternate: the current one (list[*par]) holds the members kept from
// the previous call followed by this call's parks; the last launch of a call
// appends the members that stay to the other list, and the next call's
// call_begin makes that list current.  A kept member leaves when none of its
// drifts in a call left the kepmd branch; its flag is cleared at the next
// call's start, after every fast launch of this call has read it.
struct ParkList {
    Park* buf;               // list q at buf + q * cap
    int* count2;             // entries of list 0 and list 1
    const int* par;          // index of the current list
    const int* pers;         // kept members at the start of the call (the current list's first entries)
    const int* snap;         // current-list entries when this launch's fast kernel finished
    int* leave;              // bodies leaving the slow set after this call
    int* nleave;
    unsigned char* slow;     // per body: bit 0 member, bit 1 a kept member left the kepmd branch this call,
                             // bit 2 a permanent member (cometary parameters)
    int cap;
    int chunk;               // index of this launch within the call
    int per_wave;            // general kernel: entries per 64-lane wave (spreads the latency-bound
                             // slow set over more SIMDs)
    int last;                // the call's last launch: members stay or leave
    int live = 0;            // no fast kernel runs beside this launch: the current list's count is read
                             // directly (in place of snap)
};

// Screening work the fast particle kernel defers: a box overlap with planet p
// (w = state after the first kick, then after the drift) or a body past the
// central-impact gate (w = state before the step, then after it).  The detail
// (mce_min, classification, impact time) runs in screen_requests, which keeps
// the event-building code out of the fast kernel's register allocation.
struct ScreenReq {
    int kind;      // 0 pair, 1 central
    int g;         // global body index
    int jb;        // local body index
    int step;
    int p;         // planet index (kind 0)
    int sys;
    int pad[2];
    double w[12];
};

struct ReqSink {
    ScreenReq* buf;
    int* count;
    int cap;
    int* flags;
};

__device__ __forceinline__ void push_req(const ReqSink& rs, int kind, int g, int jb, int step, int p, int sys,
                                         double a0, double a1, double a2, double a3, double a4, double a5,
                                         double b0, double b1, double b2, double b3, double b4, double b5) {
    int slot = atomicAdd(rs.count, 1);
    if (slot >= rs.cap) { atomicOr(rs.flags, FAULT_REQ_OVERFLOW); return; }
    ScreenReq* r = rs.buf + slot;
    r->kind = kind; r->g = g; r->jb = jb; r->step = step; r->p = p; r->sys = sys;
    r->w[0] = a0; r->w[1] = a1; r->w[2] = a2; r->w[3] = a3; r->w[4] = a4; r->w[5] = a5;
    r->w[6] = b0; r->w[7] = b1; r->w[8] = b2; r->w[9] = b3; r->w[10] = b4; r->w[11] = b5;
}

__device__ __forceinline__ int reserve_event(const EventSink& es) {
    int slot = atomicAdd(es.count, 1);
    if (slot >= es.cap) { atomicOr(es.flags, FAULT_EVENT_OVERFLOW); return -1; }
    return slot;
}

// ---------------------------------------------------------------------------
// mfo_obl for one body (mercury6_2.for:5191-5217): acceleration before the
// central reaction is subtracted.
template <class Ops = IeeeOps>
__device__ __forceinline__ void obl_body(double mu, double j2, double j4, double j6, double x, double y,
                                         double z, double& ax, double& ay, double& az) {
    double r2 = x * x + y * y + z * z;
    double r_1 = Ops::rcp(Ops::sqrt(r2));
    double r_2 = r_1 * r_1;
    double r_3 = r_2 * r_1;
    double jr2 = j2 * r_2;
    double jr4 = j4 * r_2 * r_2;
    double jr6 = j6 * r_2 * r_2 * r_2;
    double u2 = z * z * r_2;
    double u4 = u2 * u2;
    double u6 = u4 * u2;
    double tmp1 = mu * r_3;
    double tmp2 = jr2 * (7.5 * u2 - 1.5) + jr4 * (39.375 * u4 - 26.25 * u2 + 1.875)
                + jr6 * (187.6875 * u6 - 216.5625 * u4 + 59.0625 * u2 - 2.1875);
    double tmp3 = jr2 * 3.0 + jr4 * (17.5 * u2 - 7.5) + jr6 * (86.625 * u4 - 78.75 * u2 + 13.125);
    ax = x * tmp1 * tmp2;
    ay = y * tmp1 * tmp2;
    az = z * tmp1 * (tmp2 - tmp3);
}

// Small-body acceleration of mfo_mvs (mercury6_2.for:5034-5070):
// a = a0tp + sum_i -(m_i / d^3) (x - X_i) in planet order, plus oblateness.
// T points at one step's table (or the start table; both share the header
// and force-block layout).
// NP >= 0: the planet count is a compile-time constant (the loop unrolls and
// the table reads batch into wide scalar loads); NP < 0: runtime count np.
// Ops = RangeOps (fast kernel): 1/sqrt(s2) is exact for s2 >= 2^-767; below
// that s^-3 overflows, so both this and the IEEE form end in the non-finite
// force fault.
// The two parts of particle_accel: one planet's term of the sum (Fp: the
// planet's force block) and the header terms added after the sum.
template <class Ops, class TP>
__device__ __forceinline__ void planet_term(TP Fp, double x, double y, double z, double& a3x, double& a3y,
                                            double& a3z) {
    double dx = x - Fp[0];
    double dy = y - Fp[1];
    double dz = z - Fp[2];
    double s2 = dx * dx + dy * dy + dz * dz;
    double s_1 = Ops::rcp(Ops::sqrt(s2));
    double s_3 = s_1 * s_1 * s_1;
    double faci = Fp[3] * s_3;
    a3x = a3x - faci * dx;
    a3y = a3y - faci * dy;
    a3z = a3z - faci * dz;
}

template <class Ops, class TP>
__device__ __forceinline__ void accel_finish(const Sys& d, TP T, double x, double y, double z, double a3x,
                                             double a3y, double a3z, double& ax, double& ay, double& az) {
    ax = T[0] + a3x;
    ay = T[1] + a3y;
    az = T[2] + a3z;
    if (d.obl) {
        double ox, oy, oz;
        obl_body<Ops>(d.mu, d.j2, d.j4, d.j6, x, y, z, ox, oy, oz);
        ax = ax + (ox - T[3]);
        ay = ay + (oy - T[4]);
        az = az + (oz - T[5]);
    }
}

// TP: const double* or ConstTab* (scalar loads of the table).
template <int NP = -1, class Ops = IeeeOps, class TP>
__device__ __forceinline__ void particle_accel(const Sys& d, TP T, int np,
                                               double x, double y, double z,
                                               double& ax, double& ay, double& az) {
    const TP F = T + kHdr;
    const int npl = NP >= 0 ? NP : np;
    double a3x = 0.0, a3y = 0.0, a3z = 0.0;
#pragma unroll
    for (int p = 0; p < npl; p++) planet_term<Ops>(F + kF * p, x, y, z, a3x, a3y, a3z);
    accel_finish<Ops>(d, T, x, y, z, a3x, a3y, a3z, ax, ay, az);
}

// mce_min (mercury6_2.for:1572-1624)
__device__ inline void mce_min(double d0, double d1, double d0t, double d1t, double h,
                               double& d2min, double& tmin) {
    if (d0t * h > 0.0 || d1t * h < 0.0) {
        if (d0 <= d1) { d2min = d0; tmin = -h; }
        else { d2min = d1; tmin = 0.0; }
    } else {
        double temp = 6.0 * (d0 - d1);
        double a = temp + 3.0 * h * (d0t + d1t);
        double b = temp + 2.0 * h * (d0t + 2.0 * d1t);
        double c = h * d1t;
        temp = -.5 * (b + fsign(sqrt(fmax(b * b - 4.0 * a * c, 0.0)), b));
        double tau = (temp == 0.0) ? 0.0 : c / temp;
        tau = fmin(tau, 0.0);
        tau = fmax(tau, -1.0);
        tmin = tau * h;
        temp = 1.0 + tau;
        d2min = tau * tau * ((3.0 + 2.0 * tau) * d0 + temp * h * d0t)
              + temp * temp * ((1.0 - 2.0 * tau) * d1 + tau * h * d1t);
        d2min = fmax(d2min, 0.0);
    }
}

// One axis of mce_box (mercury6_2.for:1013-1036)
__device__ __forceinline__ void box_axis(double h, double x0, double v0, double x1, double v1,
                                         double& lo, double& hi) {
    lo = fmin(x0, x1);
    hi = fmax(x0, x1);
    if ((v0 < 0.0 && v1 > 0.0) || (v0 > 0.0 && v1 < 0.0)) {
        double temp = (v0 * x1 - v1 * x0 - .5 * h * v0 * v1) / (v0 - v1);
        lo = fmin(lo, temp);
        hi = fmax(hi, temp);
    }
}

struct Ends {   // start (after the first kick) and end (after the drift) of one body
    double x0, y0, z0, u0, v0, w0;
    double x1, y1, z1, u1, v1, w1;
};

// The pair part of mce_stat (mercury6_2.for:1820-1925) for Big body i (0-based
// local index ib) against body j (local index jb), for x-y boxes that overlap.
__device__ __noinline__ void screen_pair(const double* __restrict__ X1, const double* __restrict__ Sc, int ib,
                                         int jb, bool j_big, double mi, double mj, Ends e, double rce_j,
                                         double rphys_j, double h, double time_end, int step, int sys,
                                         EventSink es) {
    double dx0 = Sc[SP + 0] - e.x0, dy0 = Sc[SP + 1] - e.y0, dz0 = Sc[SP + 2] - e.z0;
    double du0 = Sc[SV0 + 0] - e.u0, dv0 = Sc[SV0 + 1] - e.v0, dw0 = Sc[SV0 + 2] - e.w0;
    double d0t = (dx0 * du0 + dy0 * dv0 + dz0 * dw0) * 2.0;
    double dx1 = X1[0] - e.x1, dy1 = X1[1] - e.y1, dz1 = X1[2] - e.z1;
    double du1 = Sc[SV1 + 0] - e.u1, dv1 = Sc[SV1 + 1] - e.v1, dw1 = Sc[SV1 + 2] - e.w1;
    double d1t = (dx1 * du1 + dy1 * dv1 + dz1 * dw1) * 2.0;
    double d0 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0;
    double d1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
    double d2min, tmin;
    mce_min(d0, d1, d0t, d1t, h, d2min, tmin);
    double d2ce = fmax(Sc[SRCE], rce_j);
    double d2hit = Sc[SRPH] + rphys_j;
    d2ce = d2ce * d2ce;
    d2hit = d2hit * d2hit;
    double d2near = d2hit * 4.0;
    bool clo = (d2min <= d2ce && d0t * h <= 0.0 && d1t * h >= 0.0) || (d2min <= d2hit);
    bool near = d2min <= d2near;
    if (!clo && !near) return;
    int slot = reserve_event(es);
    if (slot < 0) return;
    DevEvent* ev = es.buf + slot;
    bool swap = j_big && mj > mi;
    ev->i = (swap ? jb : ib) + 1;
    ev->j = (swap ? ib : jb) + 1;
    ev->kind = (clo ? 1 : 0) | (near ? 2 : 0) | (d2min <= d2hit ? 4 : 0);
    ev->now = d1 <= d2hit;
    ev->distance = sqrt(fmax(0.0, d2min));
    ev->time = tmin + time_end;
    ev->step = step;
    ev->sys = sys;
    double hm1 = 1.0 / h;
    double tmp0 = -tmin * hm1;
    double tmp1 = 1.0 + tmin * hm1;
    double* big = swap ? ev->xj : ev->xi;
    double* other = swap ? ev->xi : ev->xj;
    big[0] = tmp0 * Sc[SP] + tmp1 * X1[0];
    big[1] = tmp0 * Sc[SP + 1] + tmp1 * X1[1];
    big[2] = tmp0 * Sc[SP + 2] + tmp1 * X1[2];
    big[3] = tmp0 * Sc[SV0] + tmp1 * Sc[SV1];
    big[4] = tmp0 * Sc[SV0 + 1] + tmp1 * Sc[SV1 + 1];
    big[5] = tmp0 * Sc[SV0 + 2] + tmp1 * Sc[SV1 + 2];
    other[0] = tmp0 * e.x0 + tmp1 * e.x1;
    other[1] = tmp0 * e.y0 + tmp1 * e.y1;
    other[2] = tmp0 * e.z0 + tmp1 * e.z1;
    other[3] = tmp0 * e.u0 + tmp1 * e.u1;
    other[4] = tmp0 * e.v0 + tmp1 * e.v1;
    other[5] = tmp0 * e.w0 + tmp1 * e.w1;
}

// Encounter screen of body jb against the Big bodies 1..min(jb,nbig)-1 of one
// step's table (mce_stat loop order is restored by the host's sort).
// TP: const double* or ConstTab* (a table every lane of the wave shares).
template <class TP>
__device__ inline void screen_body(TP T, int np, int jb, int nbig, double mj,
                                   const Ends& e, double rce_j, double rphys_j, double h, double time_end,
                                   int step, int sys, const EventSink& es) {
    double xmin, xmax, ymin, ymax;
    box_axis(h, e.x0, e.u0, e.x1, e.u1, xmin, xmax);
    box_axis(h, e.y0, e.v0, e.y1, e.v1, ymin, ymax);
    double temp = rce_j * 1.2;
    xmin = xmin - temp; xmax = xmax + temp;
    ymin = ymin - temp; ymax = ymax + temp;
    const TP F = T + kHdr;
    const TP Sc = F + kF * np;
    int lim = jb < nbig ? jb - 1 : np;
    for (int p = 0; p < lim; p++) {
        const TP B = Sc + kScr * p;
        if ((B[SX1] >= xmin) & (xmax >= B[SX0]) & (B[SY1] >= ymin) & (ymax >= B[SY0])) {
            screen_pair((const double*)(F + kF * p), (const double*)B, p + 1, jb, jb < nbig, F[kF * p + 3], mj, e,
                        rce_j, rphys_j, h, time_end, step, sys, es);
        }
    }
}

// The part of mce_cent (mercury6_2.for:1100-1180) past its gate: two-body
// pericentre and, for an impact, the entry-time estimate.
__device__ __noinline__ void central_event(double mcen, double mj, double rcen, double h, double time_end,
                                           double xs, double ys, double zs, double us, double vs, double ws,
                                           double rr0, double rr1, double rv0, int jb, int step, int sys,
                                           EventSink es) {
    double hx = ys * ws - zs * vs;
    double hy = zs * us - xs * ws;
    double hz = xs * vs - ys * us;
    double v2 = us * us + vs * vs + ws * ws;
    double h2 = hx * hx + hy * hy + hz * hz;
    double p = h2 / (mcen + mj);
    double r0 = sqrt(rr0);
    double temp = 1.0 + p * (v2 / (mcen + mj) - 2.0 / r0);
    double e = sqrt(fmax(temp, 0.0));
    double q = p / (1.0 + e);
    if (!(q <= rcen)) return;
    temp = 0.0;
    if (e > 0.0 && e < 1.0) {
        double a = q / (1.0 - e);
        double uhit = fsign(acos(fmax(-1.0, fmin(1.0, (1.0 - rcen / a) / e))), -h);
        double u0 = fsign(acos(fmax(-1.0, fmin(1.0, (1.0 - r0 / a) / e))), rv0);
        double mhit = uhit - e * sin(uhit);
        double m0 = u0 - e * sin(u0);
        temp = (mhit - m0) / sqrt((mcen + mj) / (a * a * a));
    } else if (e > 1.0) {
        double a = q / (e - 1.0);
        double uhit = fsign(acosh(fmax(1.0, (1.0 + rcen / a) / e)), -h);
        double u0 = fsign(acosh(fmax(1.0, (1.0 + r0 / a) / e)), rv0);
        double mhit = e * sinh(uhit) - uhit;
        double m0 = e * sinh(u0) - u0;
        temp = (mhit - m0) / sqrt((mcen + mj) / (a * a * a));
    } else if (e == 1.0 && q > 0.0) {
        double uhit = fsign(sqrt(fmax(0.0, rcen / q - 1.0)), -h);
        double u0 = fsign(sqrt(fmax(0.0, r0 / q - 1.0)), rv0);
        temp = sqrt(2.0 * (q * q * q) / (mcen + mj)) * (uhit + (uhit * uhit * uhit) / 3.0 - u0
                                                        - (u0 * u0 * u0) / 3.0);
    }
    if (q == 0.0 && rr1 != rr0) temp = h * fmax(0.0, fmin(1.0, (r0 - rcen) / (r0 - sqrt(rr1))));
    int slot = reserve_event(es);
    if (slot < 0) return;
    DevEvent* ev = es.buf + slot;
    ev->i = 1;
    ev->j = jb + 1;
    ev->kind = 8;
    ev->now = 0;
    ev->distance = rcen;
    ev->time = time_end - h + temp;
    for (int k = 0; k < 6; k++) { ev->xi[k] = 0.0; ev->xj[k] = 0.0; }
    ev->step = step;
    ev->sys = sys;
}

// mce_cent (mercury6_2.for:1059-1185) for one body: (xs,vs) before the step,
// (xe,ve) after it.
__device__ __forceinline__ void central_check(double mcen, double mj, double rcen, double h, double time_end,
                                              double xs, double ys, double zs, double us, double vs, double ws,
                                              double xe, double ye, double ze, double ue, double ve, double we,
                                              int jb, int step, int sys, const EventSink& es) {
    double rcen2 = rcen * rcen;
    double rr0 = xs * xs + ys * ys + zs * zs;
    double rr1 = xe * xe + ye * ye + ze * ze;
    double rv0 = us * xs + vs * ys + ws * zs;
    double rv1 = ue * xe + ve * ye + we * ze;
    if (!((rv0 * h <= 0.0 && rv1 * h >= 0.0) || fmin(rr0, rr1) <= rcen2)) return;
    central_event(mcen, mj, rcen, h, time_end, xs, ys, zs, us, vs, ws, rr0, rr1, rv0, jb, step, sys, es);
}

// ---------------------------------------------------------------------------
// Planet prologue: one 64-lane workgroup per system, lane p owns Big body p+1.
// Start of a call, before its first planet launch (one workgroup): bodies
// that left the slow set in the previous call rejoin the fast kernel, the
// per-launch snapshots and request counts clear, the list of kept members
// becomes current, and the event count and fault flags clear.
struct CallBeginArgs {
    unsigned char* slow;
    const int* leave;
    int* count2;       // entries of list 0 and list 1
    int* par;          // current list
    int* pers;         // kept members that start the call
    int* nleave;
    int* zero;         // per-launch snapshots and request counts
    int nzero;
    int* devcount;
    int* fault;
};

__device__ __forceinline__ void call_begin_body(const CallBeginArgs& a) {
    const int nl = *a.nleave;
    for (int i = threadIdx.x; i < nl; i += blockDim.x) a.slow[a.leave[i]] = 0;
    for (int i = threadIdx.x; i < a.nzero; i += blockDim.x) a.zero[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int par = *a.par ^ 1;
        *a.par = par;
        *a.pers = a.count2[par];
        a.count2[1 - par] = 0;
        *a.nleave = 0;
        *a.devcount = 0;
        *a.fault = 0;
    }
    __syncthreads();
}

struct PlanetArgs {
    const Sys* sys;
    State st;
    double* tab;      // this launch's table block (all systems)
    double* tcur;     // per-system time, carried between launches
    EventSink es;
    double time0;     // start time when set_time != 0
    int set_time;
    int steps;
    int reset;        // compute accelerations from positions first (dtflag != 2)
    int screen;       // screen planet pairs and central impacts in-kernel
    int save;         // write step-start copies (single-step mode)
    int* zero;        // counters cleared by the launch (park and request counts), may be null
    int nzero;
    int step0;        // steps of the call done before this launch (event step numbers)
    // a call's first launch for one system starts with call_begin's work
    // (its single workgroup does it before anything else)
    int begin = 0;
    CallBeginArgs cb{};
};

// NPL: the planet count as a compile-time constant (loops unroll, so the
// independent pair terms and LDS reads overlap) or -1 (run-time count np).
template <int NPL = -1>
__device__ inline void h2j_serial(int np_rt, const double* m, const double* jdx, const double* X, const double* V,
                                  double* XJ, double* VJ) {
    // mco_h2j (mercury6_2.for:2613-2687); jdx[p] = 1/(mtot+m0) before body p is added
    const int np = NPL >= 0 ? NPL : np_rt;
    if (np < 1) return;
    for (int k = 0; k < 3; k++) { XJ[k] = X[k]; VJ[k] = V[k]; }
    double mx = m[1] * X[0], my = m[1] * X[1], mz = m[1] * X[2];
    double mu = m[1] * V[0], mv = m[1] * V[1], mw = m[1] * V[2];
#pragma unroll
    for (int p = 1; p < np; p++) {
        double temp = jdx[p];
        XJ[3 * p + 0] = X[3 * p + 0] - temp * mx;
        XJ[3 * p + 1] = X[3 * p + 1] - temp * my;
        XJ[3 * p + 2] = X[3 * p + 2] - temp * mz;
        VJ[3 * p + 0] = V[3 * p + 0] - temp * mu;
        VJ[3 * p + 1] = V[3 * p + 1] - temp * mv;
        VJ[3 * p + 2] = V[3 * p + 2] - temp * mw;
        if (p < np - 1) {
            double mb = m[p + 1];
            mx = mx + mb * X[3 * p + 0];
            my = my + mb * X[3 * p + 1];
            mz = mz + mb * X[3 * p + 2];
            mu = mu + mb * V[3 * p + 0];
            mv = mv + mb * V[3 * p + 1];
            mw = mw + mb * V[3 * p + 2];
        }
    }
}

template <int NPL = -1>
__device__ inline void j2h_serial(int np_rt, const double* jhx, const double* XJ, const double* VJ, double* X,
                                  double* V) {
    // mco_j2h (mercury6_2.for:2700-2770); jhx[p] = m_b/(mtot+m0) after body p is added
    const int np = NPL >= 0 ? NPL : np_rt;
    if (np < 1) return;
    for (int k = 0; k < 3; k++) { X[k] = XJ[k]; V[k] = VJ[k]; }
    double temp = jhx[0];
    double mx = temp * XJ[0], my = temp * XJ[1], mz = temp * XJ[2];
    double mu = temp * VJ[0], mv = temp * VJ[1], mw = temp * VJ[2];
#pragma unroll
    for (int p = 1; p < np; p++) {
        X[3 * p + 0] = XJ[3 * p + 0] + mx;
        X[3 * p + 1] = XJ[3 * p + 1] + my;
        X[3 * p + 2] = XJ[3 * p + 2] + mz;
        V[3 * p + 0] = VJ[3 * p + 0] + mu;
        V[3 * p + 1] = VJ[3 * p + 1] + mv;
        V[3 * p + 2] = VJ[3 * p + 2] + mw;
        if (p < np - 1) {
            temp = jhx[p];
            mx = mx + temp * XJ[3 * p + 0];
            my = my + temp * XJ[3 * p + 1];
            mz = mz + temp * XJ[3 * p + 2];
            mu = mu + temp * VJ[3 * p + 0];
            mv = mv + temp * VJ[3 * p + 1];
            mw = mw + temp * VJ[3 * p + 2];
        }
    }
}

// The same transforms with lane p < np computing body p+1 only: its prefix
// sums run over the bodies before it in body order, with the serial
// version's operations, so every lane's result has the serial bits.  The
// loads of a lane are independent of the other lanes' stores.
template <int NPL = -1>
__device__ inline void h2j_lanes(int np_rt, const double* m, const double* jdx, const double* X, const double* V,
                                 double* XJ, double* VJ) {
    const int np = NPL >= 0 ? NPL : np_rt;
    const int p = threadIdx.x;
    if (p >= np) return;
    if (p == 0) {
        for (int k = 0; k < 3; k++) { XJ[k] = X[k]; VJ[k] = V[k]; }
        return;
    }
    double mx = m[1] * X[0], my = m[1] * X[1], mz = m[1] * X[2];
    double mu = m[1] * V[0], mv = m[1] * V[1], mw = m[1] * V[2];
#pragma unroll
    for (int q = 1; q < (NPL >= 0 ? NPL : kMaxBig); q++) {
        if (NPL < 0 && q >= p) break;
        // with a compile-time count every body's terms load unconditionally
        // (in bounds) and only the sum is predicated, so the loads batch
        const double mb = m[q + 1];
        const double x0 = X[3 * q + 0], x1 = X[3 * q + 1], x2 = X[3 * q + 2];
        const double v0 = V[3 * q + 0], v1 = V[3 * q + 1], v2 = V[3 * q + 2];
        if (q < p) {
            mx = mx + mb * x0;
            my = my + mb * x1;
            mz = mz + mb * x2;
            mu = mu + mb * v0;
            mv = mv + mb * v1;
            mw = mw + mb * v2;
        }
    }
    double temp = jdx[p];
    XJ[3 * p + 0] = X[3 * p + 0] - temp * mx;
    XJ[3 * p + 1] = X[3 * p + 1] - temp * my;
    XJ[3 * p + 2] = X[3 * p + 2] - temp * mz;
    VJ[3 * p + 0] = V[3 * p + 0] - temp * mu;
    VJ[3 * p + 1] = V[3 * p + 1] - temp * mv;
    VJ[3 * p + 2] = V[3 * p + 2] - temp * mw;
}

template <int NPL = -1>
__device__ inline void j2h_lanes(int np_rt, const double* jhx, const double* XJ, const double* VJ, double* X,
                                 double* V) {
    const int np = NPL >= 0 ? NPL : np_rt;
    const int p = threadIdx.x;
    if (p >= np) return;
    if (p == 0) {
        for (int k = 0; k < 3; k++) { X[k] = XJ[k]; V[k] = VJ[k]; }
        return;
    }
    double temp = jhx[0];
    double mx = temp * XJ[0], my = temp * XJ[1], mz = temp * XJ[2];
    double mu = temp * VJ[0], mv = temp * VJ[1], mw = temp * VJ[2];
#pragma unroll
    for (int q = 1; q < (NPL >= 0 ? NPL : kMaxBig); q++) {
        if (NPL < 0 && q >= p) break;
        const double tq = jhx[q];
        const double x0 = XJ[3 * q + 0], x1 = XJ[3 * q + 1], x2 = XJ[3 * q + 2];
        const double v0 = VJ[3 * q + 0], v1 = VJ[3 * q + 1], v2 = VJ[3 * q + 2];
        if (q < p) {
            mx = mx + tq * x0;
            my = my + tq * x1;
            mz = mz + tq * x2;
            mu = mu + tq * v0;
            mv = mv + tq * v1;
            mw = mw + tq * v2;
        }
    }
    X[3 * p + 0] = XJ[3 * p + 0] + mx;
    X[3 * p + 1] = XJ[3 * p + 1] + my;
    X[3 * p + 2] = XJ[3 * p + 2] + mz;
    V[3 * p + 0] = VJ[3 * p + 0] + mu;
    V[3 * p + 1] = VJ[3 * p + 1] + mv;
    V[3 * p + 2] = VJ[3 * p + 2] + mw;
}

// Big-body part of mfo_mvs (mercury6_2.for:4945-5025, 5062-5070) for the
// planets in LDS.  Lane p computes body p+1; serial recurrences run in lane 0.
// Writes A (Big accelerations) and the table header terms a0tp and acen.
// NPL >= 0: the Big-Big terms are formed lane-parallel, one unordered pair
// (i < j) per lane: s^-3 once from dx = X_j - X_i (the difference the Fortran's
// pair loop forms for either body), both bodies' products into pt (LDS,
// [body][partner][3]); each body then sums its partners' products in the
// Fortran's order.  NPL < 0: each lane forms its own body's terms.
template <int NPL = -1>
__device__ inline void big_forces(const Sys& d, int np_rt, const double* m, const double* minside,
                                  const double* X, const double* XJ, double* A, double* t0, double* t2,
                                  double* a1, double* aob, double* hdr, double* pt = nullptr) {
    const int np = NPL >= 0 ? NPL : np_rt;
    int lane = threadIdx.x;
    int p = lane;
    if (NPL >= 2) {
        constexpr int NPAIR = NPL * (NPL - 1) / 2;
        for (int w = lane; w < NPAIR; w += 64) {
            int i = 0, rem = w;
            while (rem >= NPL - 1 - i) { rem -= NPL - 1 - i; i++; }
            const int j = i + 1 + rem;
            double dx = X[3 * j] - X[3 * i];
            double dy = X[3 * j + 1] - X[3 * i + 1];
            double dz = X[3 * j + 2] - X[3 * i + 2];
            double s2 = dx * dx + dy * dy + dz * dz;
            double s_1 = 1.0 / sqrt(s2);
            double s_3 = s_1 * s_1 * s_1;
            double fi = m[i + 1] * s_3;        // body j's term from partner i (before it)
            double fj = m[j + 1] * s_3;        // body i's term from partner j (after it)
            double* Tj = pt + 3 * (j * NPL + i);
            double* Ti = pt + 3 * (i * NPL + j);
            Tj[0] = fi * dx; Tj[1] = fi * dy; Tj[2] = fi * dz;
            Ti[0] = fj * dx; Ti[1] = fj * dy; Ti[2] = fj * dz;
        }
        __syncthreads();
    }
    if (p < np) {
        // a1, and the terms summed serially into a0 and a2, for Fortran k = p+2 >= 3
        if (p >= 1) {
            const double* xk = X + 3 * p;
            const double* xjk = XJ + 3 * p;
            double r2 = xk[0] * xk[0] + xk[1] * xk[1] + xk[2] * xk[2];
            double rj2 = xjk[0] * xjk[0] + xjk[1] * xjk[1] + xjk[2] * xjk[2];
            double r = 1.0 / sqrt(r2);
            double rj = 1.0 / sqrt(rj2);
            double r3 = r * r * r;
            double rj3 = rj * rj * rj;
            double mk = m[p + 1];
            double fac0 = mk * r3;
            double fac12 = m[0] * rj3;
            double fac2 = mk * fac12 / (minside[p] + m[0]);
            double q = (r2 - rj2) * .5 / rj2;
            double q2 = q * q, q3 = q * q2, q4 = q2 * q2, q5 = q2 * q3, q6 = q3 * q3, q7 = q3 * q4;
            double fac1 = 402.1875 * q7 - 187.6875 * q6 + 86.625 * q5 - 39.375 * q4 + 17.5 * q3
                        - 7.5 * q2 + 3.0 * q - 1.0;
            for (int c = 0; c < 3; c++) {
                t0[3 * p + c] = fac0 * xk[c];
                a1[3 * p + c] = fac12 * (xjk[c] + fac1 * xk[c]);
                t2[3 * p + c] = fac2 * xjk[c];
            }
        } else {
            for (int c = 0; c < 3; c++) { t0[c] = 0.0; a1[c] = 0.0; t2[c] = 0.0; }
        }
        // a3 over Big pairs, in the Fortran's pair order for this body: the
        // partners before it (subtracted), then those after it (added)
        double a3[3] = {0.0, 0.0, 0.0};
        const double* xb = X + 3 * p;
        if (NPL >= 2) {
            // the partners before this body subtract, those after it add
#pragma unroll
            for (int q = 0; q < NPL; q++) {
                const double* Tq = pt + 3 * (p * NPL + q);
                const double tx = Tq[0], ty = Tq[1], tz = Tq[2];
                if (q < p) { a3[0] = a3[0] - tx; a3[1] = a3[1] - ty; a3[2] = a3[2] - tz; }
                else if (q > p) { a3[0] = a3[0] + tx; a3[1] = a3[1] + ty; a3[2] = a3[2] + tz; }
            }
        } else if (NPL >= 0) {
            // one planet: no Big-Big pair
        } else {
            for (int i = 0; i < p; i++) {
                double dx = xb[0] - X[3 * i], dy = xb[1] - X[3 * i + 1], dz = xb[2] - X[3 * i + 2];
                double s2 = dx * dx + dy * dy + dz * dz;
                double s_1 = 1.0 / sqrt(s2);
                double s_3 = s_1 * s_1 * s_1;
                double faci = m[i + 1] * s_3;
                a3[0] = a3[0] - faci * dx; a3[1] = a3[1] - faci * dy; a3[2] = a3[2] - faci * dz;
            }
            for (int j = p + 1; j < np; j++) {
                double dx = X[3 * j] - xb[0], dy = X[3 * j + 1] - xb[1], dz = X[3 * j + 2] - xb[2];
                double s2 = dx * dx + dy * dy + dz * dz;
                double s_1 = 1.0 / sqrt(s2);
                double s_3 = s_1 * s_1 * s_1;
                double facj = m[j + 1] * s_3;
                a3[0] = a3[0] + facj * dx; a3[1] = a3[1] + facj * dy; a3[2] = a3[2] + facj * dz;
            }
        }
        for (int c = 0; c < 3; c++) A[3 * p + c] = a3[c];
        if (d.obl) obl_body(d.mu, d.j2, d.j4, d.j6, xb[0], xb[1], xb[2], aob[3 * p], aob[3 * p + 1],
                            aob[3 * p + 2]);
    }
    __syncthreads();
    // the serial sums of mfo_mvs, each lane for its own body: a0 over every
    // planet, a2 over the planets up to this one, both in body order, then
    // the body's acceleration with the Fortran's expression; lane 0 also
    // writes the particles' a0tp and the oblateness reaction
    if (lane < (np > 0 ? np : 1)) {
        double a0[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 1; q < (NPL >= 0 ? NPL : kMaxBig); q++) {
            if (q >= np) break;
            for (int c = 0; c < 3; c++) a0[c] = a0[c] - t0[3 * q + c];
        }
        double acen[3] = {0.0, 0.0, 0.0};
        if (d.obl) {
            for (int q = 0; q < np; q++) {
                double tmp4 = m[q + 1] / m[0];
                for (int c = 0; c < 3; c++) acen[c] = acen[c] - tmp4 * aob[3 * q + c];
            }
        }
        if (lane < np) {
            double a2[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 1; q < (NPL >= 0 ? NPL : kMaxBig); q++) {
                if (NPL < 0 && q > lane) break;
                const double u0 = t2[3 * q], u1 = t2[3 * q + 1], u2 = t2[3 * q + 2];
                if (q <= lane) { a2[0] = a2[0] + u0; a2[1] = a2[1] + u1; a2[2] = a2[2] + u2; }
            }
            const int q = lane;
            for (int c = 0; c < 3; c++) {
                double a = a0[c] + a1[3 * q + c] + (q >= 1 ? a2[c] : 0.0) + A[3 * q + c];
                if (d.obl) a = a + (aob[3 * q + c] - acen[c]);
                A[3 * q + c] = a;
            }
        }
        if (lane == 0) {
            double a0tp[3] = {0.0, 0.0, 0.0};
            if (np >= 1) {
                double r2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
                double r = 1.0 / sqrt(r2);
                double r3 = r * r * r;
                double fac0 = m[1] * r3;
                for (int c = 0; c < 3; c++) a0tp[c] = a0[c] - fac0 * X[c];
            }
            for (int c = 0; c < 3; c++) { hdr[c] = a0tp[c]; hdr[3 + c] = acen[c]; }
        }
    }
    __syncthreads();
}

// Encounter screen of Big body jb against the Big bodies before it (the
// planet kernel's part of mce_stat): screen_body with the planet loop
// unrolled when the planet count is a compile-time constant.
// box: the planets' padded x-y boxes of this step (LDS copy of the table's).
template <int NPL>
__device__ inline void screen_planet(const double* __restrict__ T, const double* box, int np_rt, int jb, double mj,
                                     const Ends& e, double rce_j, double rphys_j, double h, double time_end, int step,
                                     int sys, const EventSink& es) {
    if (NPL < 0) {
        screen_body(T, np_rt, jb, np_rt + 1, mj, e, rce_j, rphys_j, h, time_end, step, sys, es);
        return;
    }
    double xmin, xmax, ymin, ymax;
    box_axis(h, e.x0, e.u0, e.x1, e.u1, xmin, xmax);
    box_axis(h, e.y0, e.v0, e.y1, e.v1, ymin, ymax);
    double temp = rce_j * 1.2;
    xmin = xmin - temp; xmax = xmax + temp;
    ymin = ymin - temp; ymax = ymax + temp;
    const double* F = T + kHdr;
    const double* Sc = F + kF * NPL;
#pragma unroll
    for (int p = 0; p < NPL; p++) {
        const double* b = box + 4 * p;
        if (p < jb - 1 && b[1] >= xmin && xmax >= b[0] && b[3] >= ymin && ymax >= b[2])
            screen_pair(F + kF * p, Sc + kScr * p, p + 1, jb, true, F[kF * p + 3], mj, e, rce_j, rphys_j, h,
                        time_end, step, sys, es);
    }
}

// One wave per system, latency-bound; raised issue priority keeps it ahead of
// the particle waves it shares a SIMD with.
struct PlanetLds {
    double m[kMaxBig], minside[kMaxBig], jdx[kMaxBig], jhx[kMaxBig], gm[kMaxBig];
    double X[3 * kMaxBig], V[3 * kMaxBig], A[3 * kMaxBig], XJ[3 * kMaxBig], VJ[3 * kMaxBig];
    double X0[3 * kMaxBig], V0[3 * kMaxBig], t0[3 * kMaxBig], t2[3 * kMaxBig];
    double a1[3 * kMaxBig], aob[3 * kMaxBig], hdr[8];
    double box[4 * kMaxBig];   // this step's padded x-y boxes (screening)
    double pt[3 * 12 * 12];    // Big-Big pair products (big_forces, up to 12 planets unrolled)
};

template <int NPL>
__device__ __forceinline__ void planets_run(const PlanetArgs& args, PlanetLds& L) {
    const int sid = blockIdx.x;
    const Sys d = args.sys[sid];
    const int np = NPL >= 0 ? NPL : d.nbig - 1;
    const int lane = threadIdx.x;
    double* m = L.m; double* minside = L.minside; double* jdx = L.jdx; double* jhx = L.jhx; double* gm = L.gm;
    double* X = L.X; double* V = L.V; double* A = L.A; double* XJ = L.XJ; double* VJ = L.VJ;
    double* X0 = L.X0; double* V0 = L.V0; double* t0 = L.t0; double* t2 = L.t2;
    double* a1 = L.a1; double* aob = L.aob; double* hdr = L.hdr;
    const long long cap = args.st.cap;
    const int b0 = d.base;
    for (int q = lane; q < d.nbig; q += 64) m[q] = args.st.m[b0 + q];
    for (int p = lane; p < np; p += 64)
        for (int c = 0; c < 3; c++) {
            X[3 * p + c] = args.st.x[c * cap + b0 + 1 + p];
            V[3 * p + c] = args.st.v[c * cap + b0 + 1 + p];
            A[3 * p + c] = args.st.a[c * cap + b0 + 1 + p];
        }
    // this lane's planet: radii read once for the launch
    double rce_p = 0.0, rphys_p = 0.0;
    if (lane < np) { rce_p = args.st.rce[b0 + 1 + lane]; rphys_p = args.st.rphys[b0 + 1 + lane]; }
    __syncthreads();
    if (lane == 0) {
        // constants of the Jacobi transforms and drifts (mdt_mvs mercury6_2.for:4041-4046)
        double ms = 0.0;
        for (int p = 0; p < np; p++) { minside[p] = ms; ms = ms + m[p + 1]; }
        double mtot = (np >= 1) ? m[1] : 0.0;
        for (int p = 1; p < np; p++) { jdx[p] = 1.0 / (mtot + m[0]); mtot = mtot + m[p + 1]; }
        mtot = (np >= 1) ? m[1] : 0.0;
        if (np >= 1) jhx[0] = m[1] / (mtot + m[0]);
        for (int p = 1; p < np; p++) { mtot = mtot + m[p + 1]; jhx[p] = m[p + 1] / (mtot + m[0]); }
        double mins = m[0];
        for (int p = 0; p < np; p++) {
            double msofar = mins + m[p + 1];
            gm[p] = m[0] * msofar / mins;
            mins = msofar;
        }
    }
    __syncthreads();
    const int ts = d.ts;
    if (args.reset) {
        if (lane == 0) h2j_serial<NPL>(np, m, jdx, X, V, XJ, VJ);
        __syncthreads();
        big_forces<NPL>(d, np, m, minside, X, XJ, A, t0, t2, a1, aob, hdr, L.pt);
        if (lane < np && !(isfinite(A[3 * lane]) && isfinite(A[3 * lane + 1]) && isfinite(A[3 * lane + 2])))
            atomicOr(args.st.fault, FAULT_FORCE);
        double* St = args.tab + d.start;
        if (lane < 8) St[lane] = (lane < 6) ? hdr[lane] : 0.0;
        for (int p = lane; p < np; p += 64) {
            for (int c = 0; c < 3; c++) St[kHdr + kF * p + c] = X[3 * p + c];
            St[kHdr + kF * p + 3] = m[p + 1];
        }
    }
    double time_end = args.set_time ? args.time0 : args.tcur[sid];
    if (args.zero && sid == 0 && lane < args.nzero) args.zero[lane] = 0;
    const double h = d.h, hby2 = d.hby2;
    for (int s = 0; s < args.steps; s++) {
        double* T = args.tab + d.tab + (long long)s * ts;
        time_end = time_end + h;
        double xs[3], vs[3];
        int p = lane;
        if (p < np) {
            for (int c = 0; c < 3; c++) { xs[c] = X[3 * p + c]; vs[c] = V[3 * p + c]; }
            if (args.save)
                for (int c = 0; c < 3; c++) {
                    args.st.ox[c * cap + b0 + 1 + p] = xs[c];
                    args.st.ov[c * cap + b0 + 1 + p] = vs[c];
                }
            for (int c = 0; c < 3; c++) {
                V[3 * p + c] = V[3 * p + c] + hby2 * A[3 * p + c];
                X0[3 * p + c] = X[3 * p + c];
                V0[3 * p + c] = V[3 * p + c];
            }
        }
        __syncthreads();
        h2j_lanes<NPL>(np, m, jdx, X, V, XJ, VJ);
        __syncthreads();
        if (p < np) {
            Orbit o{XJ[3 * p], XJ[3 * p + 1], XJ[3 * p + 2], VJ[3 * p], VJ[3 * p + 1], VJ[3 * p + 2]};
            // the kepmd branch through drift_fast (same bits, shorter divides
            // and roots) where its gate admits the drift, else the whole chain
            const double mu = gm[p];
            const bool gate = mu >= 0x1p-100 && mu <= 0x1p100 && fabs(h) >= 0x1p-100 && fabs(h) <= 0x1p100;
            if (!(gate && drift_fast(mu, o, h, o.x * o.x + o.y * o.y + o.z * o.z))) drift_one(mu, o, h);
            XJ[3 * p] = o.x; XJ[3 * p + 1] = o.y; XJ[3 * p + 2] = o.z;
            VJ[3 * p] = o.vx; VJ[3 * p + 1] = o.vy; VJ[3 * p + 2] = o.vz;
        }
        __syncthreads();
        j2h_lanes<NPL>(np, jhx, XJ, VJ, X, V);
        __syncthreads();
        // screening data of this step (velocities after the drift, before the kick)
        double vd[3] = {0.0, 0.0, 0.0};
        if (p < np) {
            double* B = T + kHdr + kF * np + kScr * p;
            double lo, hi;
            box_axis(h, X0[3 * p], V0[3 * p], X[3 * p], V[3 * p], lo, hi);
            double pad = rce_p * 1.2;
            B[SX0] = lo - pad; B[SX1] = hi + pad;
            L.box[4 * p] = lo - pad; L.box[4 * p + 1] = hi + pad;
            box_axis(h, X0[3 * p + 1], V0[3 * p + 1], X[3 * p + 1], V[3 * p + 1], lo, hi);
            B[SY0] = lo - pad; B[SY1] = hi + pad;
            L.box[4 * p + 2] = lo - pad; L.box[4 * p + 3] = hi + pad;
            for (int c = 0; c < 3; c++) vd[c] = V[3 * p + c];
            for (int c = 0; c < 3; c++) {
                B[SP + c] = X0[3 * p + c];
                B[SV0 + c] = V0[3 * p + c];
                B[SV1 + c] = V[3 * p + c];
                T[kHdr + kF * p + c] = X[3 * p + c];
            }
            B[SRCE] = rce_p;
            B[SRPH] = rphys_p;
            T[kHdr + kF * p + 3] = m[p + 1];
            if (args.save)
                for (int c = 0; c < 3; c++) {
                    args.st.v0k[c * cap + b0 + 1 + p] = V0[3 * p + c];
                    args.st.v1[c * cap + b0 + 1 + p] = V[3 * p + c];
                }
        }
        big_forces<NPL>(d, np, m, minside, X, XJ, A, t0, t2, a1, aob, hdr, L.pt);
        if (lane < np && !(isfinite(A[3 * lane]) && isfinite(A[3 * lane + 1]) && isfinite(A[3 * lane + 2])))
            atomicOr(args.st.fault, FAULT_FORCE);
        if (lane < 8) T[lane] = (lane < 6) ? hdr[lane] : (lane == 6 ? time_end : 0.0);
        if (p < np) {
            for (int c = 0; c < 3; c++) V[3 * p + c] = V[3 * p + c] + hby2 * A[3 * p + c];
        }
        __syncthreads();
        if (args.screen && p < np) {
            // make this step's table visible to the screening lanes
            __threadfence_block();
            int jb = p + 1;
            Ends e{X0[3 * p], X0[3 * p + 1], X0[3 * p + 2], V0[3 * p], V0[3 * p + 1], V0[3 * p + 2],
                   X[3 * p], X[3 * p + 1], X[3 * p + 2], vd[0], vd[1], vd[2]};
            screen_planet<NPL>(T, L.box, np, jb, m[jb], e, rce_p, rphys_p, h, time_end, s + args.step0, sid,
                               args.es);
            central_check(m[0], m[jb], d.rcen, h, time_end, xs[0], xs[1], xs[2], vs[0], vs[1], vs[2],
                          X[3 * p], X[3 * p + 1], X[3 * p + 2], V[3 * p], V[3 * p + 1], V[3 * p + 2], jb,
                          s + args.step0, sid, args.es);
        }
        __syncthreads();
    }
    for (int p = lane; p < np; p += 64)
        for (int c = 0; c < 3; c++) {
            args.st.x[c * cap + b0 + 1 + p] = X[3 * p + c];
            args.st.v[c * cap + b0 + 1 + p] = V[3 * p + c];
            args.st.a[c * cap + b0 + 1 + p] = A[3 * p + c];
        }
    if (lane == 0) args.tcur[sid] = time_end;
}

// The planet kernel: one 64-lane workgroup per system, lane p owns Big body
// p+1.  Each workgroup branches once on its system's planet count into the
// body built for it (0..12); more planets take the run-time-count body.
__global__ void __launch_bounds__(64) mvs_planets(PlanetArgs args) {
    __builtin_amdgcn_s_setprio(3);
    __shared__ PlanetLds L;
    if (args.begin) call_begin_body(args.cb);
    switch (args.sys[blockIdx.x].nbig - 1) {
        case 0: planets_run<0>(args, L); break;
        case 1: planets_run<1>(args, L); break;
        case 2: planets_run<2>(args, L); break;
        case 3: planets_run<3>(args, L); break;
        case 4: planets_run<4>(args, L); break;
        case 5: planets_run<5>(args, L); break;
        case 6: planets_run<6>(args, L); break;
        case 7: planets_run<7>(args, L); break;
        case 8: planets_run<8>(args, L); break;
        case 9: planets_run<9>(args, L); break;
        case 10: planets_run<10>(args, L); break;
        case 11: planets_run<11>(args, L); break;
        case 12: planets_run<12>(args, L); break;
        default: planets_run<-1>(args, L); break;
    }
}

// ---------------------------------------------------------------------------
// Symplectic corrector mco_mvs2h (mercury6_2.for:2107-2219) on a copy of the
// state: for k = 1, 2: drift ha(k), kick -hb(k), drift -2 ha(k), kick +hb(k),
// drift ha(k).  The planet wave writes the four force points' tables (header
// and force block) and the corrected planet states; the particle kernel then
// applies the same sequence to every particle.
struct CorrectorArgs {
    const Sys* sys;
    State st;
    double* tab;       // 4 tables of stride Sys::ts
    double* xo;        // corrected positions [3][cap]
    double* vo;        // corrected velocities [3][cap]
    double ha[2], hb[2];
};

__global__ void __launch_bounds__(64) corrector_planets(CorrectorArgs args) {
    const Sys d = args.sys[0];
    const int np = d.nbig - 1;
    const int lane = threadIdx.x;
    __shared__ double m[kMaxBig], minside[kMaxBig], jdx[kMaxBig], jhx[kMaxBig], gm[kMaxBig];
    __shared__ double X[3 * kMaxBig], V[3 * kMaxBig], A[3 * kMaxBig], XJ[3 * kMaxBig], VJ[3 * kMaxBig];
    __shared__ double t0[3 * kMaxBig], t2[3 * kMaxBig], a1[3 * kMaxBig], aob[3 * kMaxBig], hdr[8];
    const long long cap = args.st.cap;
    const int b0 = d.base;
    for (int q = lane; q < d.nbig; q += 64) m[q] = args.st.m[b0 + q];
    for (int p = lane; p < np; p += 64)
        for (int c = 0; c < 3; c++) {
            X[3 * p + c] = args.st.x[c * cap + b0 + 1 + p];
            V[3 * p + c] = args.st.v[c * cap + b0 + 1 + p];
        }
    __syncthreads();
    if (lane == 0) {
        double ms = 0.0;
        for (int p = 0; p < np; p++) { minside[p] = ms; ms = ms + m[p + 1]; }
        double mtot = (np >= 1) ? m[1] : 0.0;
        for (int p = 1; p < np; p++) { jdx[p] = 1.0 / (mtot + m[0]); mtot = mtot + m[p + 1]; }
        mtot = (np >= 1) ? m[1] : 0.0;
        if (np >= 1) jhx[0] = m[1] / (mtot + m[0]);
        for (int p = 1; p < np; p++) { mtot = mtot + m[p + 1]; jhx[p] = m[p + 1] / (mtot + m[0]); }
        double mins = m[0];
        for (int p = 0; p < np; p++) {
            double msofar = mins + m[p + 1];
            gm[p] = m[0] * msofar / mins;
            mins = msofar;
        }
    }
    __syncthreads();
    int fp = 0;   // force point 0..3
    for (int k = 0; k < 2; k++) {
        const double dts[3] = {args.ha[k], -2.0 * args.ha[k], args.ha[k]};
        for (int phase = 0; phase < 3; phase++) {
            if (lane == 0) h2j_serial(np, m, jdx, X, V, XJ, VJ);
            __syncthreads();
            if (lane < np) {
                Orbit o{XJ[3 * lane], XJ[3 * lane + 1], XJ[3 * lane + 2], VJ[3 * lane], VJ[3 * lane + 1],
                        VJ[3 * lane + 2]};
                drift_one(gm[lane], o, dts[phase]);
                XJ[3 * lane] = o.x; XJ[3 * lane + 1] = o.y; XJ[3 * lane + 2] = o.z;
                VJ[3 * lane] = o.vx; VJ[3 * lane + 1] = o.vy; VJ[3 * lane + 2] = o.vz;
            }
            __syncthreads();
            if (lane == 0) j2h_serial(np, jhx, XJ, VJ, X, V);
            __syncthreads();
            if (phase == 2) break;
            big_forces(d, np, m, minside, X, XJ, A, t0, t2, a1, aob, hdr);
            double* T = args.tab + (long long)fp * d.ts;
            if (lane < 8) T[lane] = (lane < 6) ? hdr[lane] : 0.0;
            for (int p = lane; p < np; p += 64) {
                for (int c = 0; c < 3; c++) T[kHdr + kF * p + c] = X[3 * p + c];
                T[kHdr + kF * p + 3] = m[p + 1];
            }
            if (lane < np)
                for (int c = 0; c < 3; c++) {
                    if (phase == 0) V[3 * lane + c] = V[3 * lane + c] - args.hb[k] * A[3 * lane + c];
                    else V[3 * lane + c] = V[3 * lane + c] + args.hb[k] * A[3 * lane + c];
                }
            __syncthreads();
            fp++;
        }
    }
    for (int p = lane; p < np; p += 64)
        for (int c = 0; c < 3; c++) {
            args.xo[c * cap + b0 + 1 + p] = X[3 * p + c];
            args.vo[c * cap + b0 + 1 + p] = V[3 * p + c];
        }
    if (lane < 3) { args.xo[lane * cap + b0] = 0.0; args.vo[lane * cap + b0] = 0.0; }
}

// NG (cometary model): each kick adds mfo_ngf's angf at the kick point, as
// hb * (angf + ausr + a) with no user force (mercury6_2.for:2171-2176,
// 2200-2205).
template <bool NG>
__global__ void __launch_bounds__(128) corrector_particles(CorrectorArgs args) {
    const Sys d = args.sys[0];
    const int jb = d.nbig + blockIdx.x * blockDim.x + threadIdx.x;
    if (jb >= d.n) return;
    const long long cap = args.st.cap;
    const int g = d.base + jb;
    const int np = d.nbig - 1;
    Orbit o{args.st.x[g], args.st.x[cap + g], args.st.x[2 * cap + g],
            args.st.v[g], args.st.v[cap + g], args.st.v[2 * cap + g]};
    double P[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (NG) load_ngf(args.st, g, P);
    double n0 = 0.0, n1 = 0.0, n2 = 0.0;
    int fp = 0;
    for (int k = 0; k < 2; k++) {
        drift_one(d.mu, o, args.ha[k]);
        double ax, ay, az;
        particle_accel(d, args.tab + (long long)fp * d.ts, np, o.x, o.y, o.z, ax, ay, az);
        if (NG) {
            ngf_body(P[0], P[1], P[2], P[4], o.x, o.y, o.z, o.vx, o.vy, o.vz, n0, n1, n2, args.st.fault);
            o.vx = o.vx - args.hb[k] * ((n0 + 0.0) + ax);
            o.vy = o.vy - args.hb[k] * ((n1 + 0.0) + ay);
            o.vz = o.vz - args.hb[k] * ((n2 + 0.0) + az);
        } else {
            o.vx = o.vx - args.hb[k] * ax; o.vy = o.vy - args.hb[k] * ay; o.vz = o.vz - args.hb[k] * az;
        }
        fp++;
        drift_one(d.mu, o, -2.0 * args.ha[k]);
        particle_accel(d, args.tab + (long long)fp * d.ts, np, o.x, o.y, o.z, ax, ay, az);
        if (NG) {
            ngf_body(P[0], P[1], P[2], P[4], o.x, o.y, o.z, o.vx, o.vy, o.vz, n0, n1, n2, args.st.fault);
            o.vx = o.vx + args.hb[k] * ((n0 + 0.0) + ax);
            o.vy = o.vy + args.hb[k] * ((n1 + 0.0) + ay);
            o.vz = o.vz + args.hb[k] * ((n2 + 0.0) + az);
        } else {
            o.vx = o.vx + args.hb[k] * ax; o.vy = o.vy + args.hb[k] * ay; o.vz = o.vz + args.hb[k] * az;
        }
        fp++;
        drift_one(d.mu, o, args.ha[k]);
    }
    args.xo[g] = o.x; args.xo[cap + g] = o.y; args.xo[2 * cap + g] = o.z;
    args.vo[g] = o.vx; args.vo[cap + g] = o.vy; args.vo[2 * cap + g] = o.vz;
}

// ---------------------------------------------------------------------------
// Particle passes.
struct ParticleArgs {
    const Sys* sys;
    const int2* tiles;   // (system, first particle index) per workgroup
    State st;
    const double* tab;
    EventSink es;
    ParkList park;
    ReqSink req;
    unsigned long long* work;   // [1] particle-steps done by the general kernel (timing mode; the
                                // fast kernel's are the call's total less these)
    int steps;
    int reset;
    int screen;
    int save;
    int step0;
    // general kernel: the launch runs the steps of its block for the slow set,
    // after the rest of the previous launch's block for the members that
    // launch's fast kernel parked (ptab: that block, psteps: its step count;
    // 0 for the call's first launch)
    const double* ptab = nullptr;
    int psteps = 0;
    // fast kernel: particles per lane; lane t of a tile takes its particles
    // t, t + blockDim.x, ... (ppl of them) one after the other
    int ppl = 1;
};

// Adds each lane's completed particle-steps into *counter, one atomic per wave.
__device__ __forceinline__ void count_work(unsigned long long* counter, int steps) {
    int v = steps;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0 && v) atomicAdd(counter, (unsigned long long)v);
}

// Step-start copies for the host-driven event call (single-step mode).
__device__ __forceinline__ void save_step(const State& st, int g, const double xs[6], const Orbit& k0,
                                          const Orbit& o) {
    const long long cap = st.cap;
    st.ox[g] = xs[0]; st.ox[cap + g] = xs[1]; st.ox[2 * cap + g] = xs[2];
    st.ov[g] = xs[3]; st.ov[cap + g] = xs[4]; st.ov[2 * cap + g] = xs[5];
    st.v0k[g] = k0.vx; st.v0k[cap + g] = k0.vy; st.v0k[2 * cap + g] = k0.vz;
    st.v1[g] = o.vx; st.v1[cap + g] = o.vy; st.v1[2 * cap + g] = o.vz;
}

// Rest of one step after the drift: screen, force, second kick, central check.
// A non-finite acceleration raises the fault flag (the step then fails, as
// mvs_acceleration's check does in the reference, mercury_cuda_symplectic.cuh:78).
// As finish_step, with the screening details deferred to screen_requests:
// inline are only the x-y box tests of mce_stat and the gate of mce_cent.
//
// The states a request records stay out of registers: E (the lane's slots in
// LDS, stride `ls`) holds the state after the first kick (E[0..5]) and the
// step-start velocities (E[6..8]; the step-start position is E[0..2]); the
// central gate needs only rr = |x|^2 and rv = v.x of the step start.
constexpr int kEnds = 9;       // doubles per lane in the fast kernel's LDS

// The box test of every planet is evaluated without branches (the four
// comparisons of mce_stat, mercury6_2.for:1810-1811, all made) into the
// wave's lane mask `hits`: each comparison's lane mask comes straight from its
// v_cmp, the masks are ANDed and ORed on the scalar unit, and an empty asm
// statement at each planet's end makes the OR happen there, so the planet's
// box bounds die in its iteration (left to the compiler, the comparisons sink
// to the mask's use after the loop and every planet's bounds stay live in
// SGPRs, which then spill into VGPR lanes).  The rare lanes with a hit then
// repeat the tests and push one request per overlapping planet in planet
// order.
template <bool SCREEN, int NP>
__device__ __forceinline__ void finish_step_deferred(const Sys& d, ConstTab* T, int np, int g,
                                                     int jb, double pad, int s, int sid, const ReqSink& rs,
                                                     int* fault, double& rr, double& rv, const double* E, int ls,
                                                     Orbit& o, double& ax, double& ay, double& az) {
    double xmin = 0.0, xmax = 0.0, ymin = 0.0, ymax = 0.0;
    if (SCREEN) {
        box_axis(d.h, E[0], E[3 * ls], o.x, o.vx, xmin, xmax);
        box_axis(d.h, E[ls], E[4 * ls], o.y, o.vy, ymin, ymax);
        xmin = xmin - pad; xmax = xmax + pad;
        ymin = ymin - pad; ymax = ymax + pad;
    }
    // per planet: the box test and the force term (each planet's table block
    // is live for one iteration); the acceleration sum stays in planet order
    const int npl = NP >= 0 ? NP : np;
    ConstTab* F = T + kHdr;
    ConstTab* Sc = F + kF * npl;
    unsigned long long hits = 0ull;
    double a3x = 0.0, a3y = 0.0, a3z = 0.0;
#pragma unroll
    for (int p = 0; p < npl; p++) {
        unsigned long long hp = 0ull;
        if (SCREEN) {
            // particles (jb >= nbig) test every planet; this kernel never sees Big bodies
            constexpr int OGE = 3;   // ordered >=: false on a NaN, as the C comparison
            ConstTab* B = Sc + kScr * p;
            hp = __builtin_amdgcn_fcmp(B[SX1], xmin, OGE) & __builtin_amdgcn_fcmp(xmax, B[SX0], OGE) &
                 __builtin_amdgcn_fcmp(B[SY1], ymin, OGE) & __builtin_amdgcn_fcmp(ymax, B[SY0], OGE);
        }
        planet_term<RangeOps>(F + kF * p, o.x, o.y, o.z, a3x, a3y, a3z);
        if (SCREEN) {
            hits |= hp;
            asm volatile("" : "+s"(hits));
        }
        // the scheduler keeps each planet's scalar loads in its own iteration
        // (hoisting every planet's 16 SGPRs of table data spills SGPRs)
        __builtin_amdgcn_sched_barrier(0);
    }
    accel_finish<RangeOps>(d, T, o.x, o.y, o.z, a3x, a3y, a3z, ax, ay, az);
    if (SCREEN) {
        if ((hits >> __lane_id()) & 1ull) {
#pragma unroll 1
            for (int p = 0; p < npl; p++) {
                ConstTab* B = Sc + kScr * p;
                if ((B[SX1] >= xmin) & (xmax >= B[SX0]) & (B[SY1] >= ymin) & (ymax >= B[SY0]))
                    push_req(rs, 0, g, jb, s, p, sid, E[0], E[ls], E[2 * ls], E[3 * ls], E[4 * ls], E[5 * ls],
                             o.x, o.y, o.z, o.vx, o.vy, o.vz);
            }
        }
    }
    if (!isfinite(ax) || !isfinite(ay) || !isfinite(az)) atomicOr(fault, FAULT_FORCE);
    o.vx = o.vx + d.hby2 * ax;
    o.vy = o.vy + d.hby2 * ay;
    o.vz = o.vz + d.hby2 * az;
    // x.x and v.x at the step's end: the next step's start values
    const double rr1 = o.x * o.x + o.y * o.y + o.z * o.z;
    if (SCREEN) {
        double rcen2 = d.rcen * d.rcen;
        double rv1 = o.vx * o.x + o.vy * o.y + o.vz * o.z;
        if ((rv * d.h <= 0.0 && rv1 * d.h >= 0.0) || fmin(rr, rr1) <= rcen2)
            push_req(rs, 1, g, jb, s, 0, sid, E[0], E[ls], E[2 * ls], E[6 * ls], E[7 * ls], E[8 * ls],
                     o.x, o.y, o.z, o.vx, o.vy, o.vz);
        rv = rv1;
    }
    rr = rr1;
}

// NG: the kick adds the cometary acceleration ang (mdt_mvs's angf, evaluated
// after the drift) as hby2 * (angf + ausr + a) with no user force.
template <bool SCREEN, bool NG = false, class TP>
__device__ __forceinline__ void finish_step(const Sys& d, TP T, int np, int jb,
                                            double mj, double rce_j, double rphys_j, int s, int sid,
                                            const EventSink& es, int* fault, const double xs[6],
                                            const Orbit& k0, Orbit& o, double& ax, double& ay, double& az,
                                            const double* ang = nullptr) {
    if (SCREEN) {
        Ends e{k0.x, k0.y, k0.z, k0.vx, k0.vy, k0.vz, o.x, o.y, o.z, o.vx, o.vy, o.vz};
        screen_body(T, np, jb, d.nbig, mj, e, rce_j, rphys_j, d.h, T[6], s, sid, es);
    }
    particle_accel(d, T, np, o.x, o.y, o.z, ax, ay, az);
    if (!isfinite(ax) || !isfinite(ay) || !isfinite(az)) atomicOr(fault, FAULT_FORCE);
    if (NG) {
        o.vx = o.vx + d.hby2 * ((ang[0] + 0.0) + ax);
        o.vy = o.vy + d.hby2 * ((ang[1] + 0.0) + ay);
        o.vz = o.vz + d.hby2 * ((ang[2] + 0.0) + az);
    } else {
        o.vx = o.vx + d.hby2 * ax;
        o.vy = o.vy + d.hby2 * ay;
        o.vz = o.vz + d.hby2 * az;
    }
    if (SCREEN)
        central_check(d.mu, mj, d.rcen, d.h, T[6], xs[0], xs[1], xs[2], xs[3], xs[4], xs[5], o.x, o.y, o.z,
                      o.vx, o.vy, o.vz, jb, s, sid, es);
}

// One particle's launch: S steps of particle g.  NP >= 0: the planet count as a
// compile-time constant (the force and box loops unroll and the table reads
// batch into wide scalar loads); NP < 0: read at run time.
template <bool SCREEN, bool SAVE, int NP>
__device__ __forceinline__ void fast_particle(const ParticleArgs& args, double* lds_ends, int sid, const Sys& d,
                                              int g) {
    const int jb = g - d.base;                     // local body index
    // lanes past the system's end, and slow-set members, idle to the end
    const bool live = jb < d.n && !args.park.slow[g];
    const long long cap = args.st.cap;
    const int np = d.nbig - 1;
    Orbit o{0, 0, 0, 0, 0, 0};
    double ax = 0, ay = 0, az = 0;
    double rce_j = 0.0;
    if (live) {
        o = Orbit{args.st.x[g], args.st.x[cap + g], args.st.x[2 * cap + g],
                  args.st.v[g], args.st.v[cap + g], args.st.v[2 * cap + g]};
        if (args.reset) {
            particle_accel<NP, RangeOps>(d, const_tab(args.tab + d.start), np, o.x, o.y, o.z, ax, ay, az);
            if (!isfinite(ax) || !isfinite(ay) || !isfinite(az)) atomicOr(args.st.fault, FAULT_FORCE);
        } else {
            ax = args.st.a[g]; ay = args.st.a[cap + g]; az = args.st.a[2 * cap + g];
        }
        if (SCREEN) rce_j = args.st.rce[g];
    }
    const double pad = rce_j * 1.2;                 // mce_stat's box padding
    const int ls = blockDim.x;
    double* E = lds_ends + threadIdx.x;
    ConstTab* tab = const_tab(args.tab + d.tab);
    const int nsteps = live ? args.steps : 0;
    // drift_fast's exact range needs mu and |h| in [2^-100, 2^100]; a system
    // outside it parks every particle at its first step
    const bool in_range = d.mu >= 0x1p-100 && d.mu <= 0x1p100 && fabs(d.h) >= 0x1p-100 && fabs(d.h) <= 0x1p100;
    // x.x and v.x of the state at each step's start (the central gate's r0^2
    // and v.x, the drift's r0^2), carried from the previous step's end
    double rr = o.x * o.x + o.y * o.y + o.z * o.z;
    double rv = SCREEN ? o.vx * o.x + o.vy * o.y + o.vz * o.z : 0.0;
    int s = 0;
    for (; s < nsteps; s++) {
        ConstTab* T = tab + (long long)s * d.ts;
        double xs[6] = {o.x, o.y, o.z, o.vx, o.vy, o.vz};
        // With screening, the step-start state also waits in LDS (E[0..2] and
        // E[6..8]), so the kick and drift update o in place and a parking lane
        // reloads it from there: no register copy of the state is kept across
        // the drift's exits.  Without screening, w is the working copy.
        Orbit w = o;
        w.vx = w.vx + d.hby2 * ax;
        w.vy = w.vy + d.hby2 * ay;
        w.vz = w.vz + d.hby2 * az;
        Orbit k0 = w;
        if (SCREEN) {
            E[0] = w.x; E[ls] = w.y; E[2 * ls] = w.z;
            E[3 * ls] = w.vx; E[4 * ls] = w.vy; E[5 * ls] = w.vz;
            E[6 * ls] = o.vx; E[7 * ls] = o.vy; E[8 * ls] = o.vz;
            o = w;
        }
        bool ok = in_range && drift_fast(d.mu, SCREEN ? o : w, d.h, rr);
        // wave-ballot compaction: lanes whose drift leaves the kepmd branch
        // park here and finish in the general kernel
        unsigned long long mask = __ballot(!ok);
        if (!ok) {
            if (SCREEN) o = Orbit{E[0], E[ls], E[2 * ls], E[6 * ls], E[7 * ls], E[8 * ls]};
            int leader = __ffsll((long long)mask) - 1;
            int lane = __lane_id();
            const int par = *args.park.par;
            int base = 0;
            if (lane == leader) base = atomicAdd(args.park.count2 + par, __popcll(mask));
            base = __shfl(base, leader);
            int slot = base + __popcll(mask & ((1ull << lane) - 1ull));
            if (slot < args.park.cap) args.park.buf[(long long)par * args.park.cap + slot] = Park{g, s, sid, args.park.chunk};
            args.park.slow[g] = 1;
            break;
        }
        if (!SCREEN) o = w;
        if (SAVE) save_step(args.st, g, xs, k0, o);
        finish_step_deferred<SCREEN, NP>(d, T, np, g, jb, pad, s, sid, args.req, args.st.fault, rr, rv, E, ls,
                                         o, ax, ay, az);
    }
    if (!live) return;
    // parked lanes store their pre-step state and step-start acceleration
    args.st.x[g] = o.x; args.st.x[cap + g] = o.y; args.st.x[2 * cap + g] = o.z;
    args.st.v[g] = o.vx; args.st.v[cap + g] = o.vy; args.st.v[2 * cap + g] = o.vz;
    args.st.a[g] = ax; args.st.a[cap + g] = ay; args.st.a[2 * cap + g] = az;
}

// One workgroup's tile: ppl particles per lane, one after the other (g is the
// global body index; lanes past the system's end idle).
template <bool SCREEN, bool SAVE, int NP>
__device__ __forceinline__ void fast_tile(const ParticleArgs& args, double* lds_ends) {
    const int2 tile = args.tiles[blockIdx.x];
    const int sid = tile.x;
    const Sys d = args.sys[sid];
#pragma unroll 1
    for (int k = 0; k < args.ppl; k++)
        fast_particle<SCREEN, SAVE, NP>(args, lds_ends, sid, d, tile.y + threadIdx.x + k * (int)blockDim.x);
}

// The fast particle kernel.  A tile never spans two systems, so each
// workgroup branches once, on its system's planet count, into the body built
// for that count: single systems and ensembles of mixed planet counts run the
// same specialised code.  SAVE (single-step mode) keeps the run-time count.
// MINW: waves per SIMD the register allocation is capped for.
// SCREEN: launched with kEnds * blockDim.x doubles of dynamic LDS.
template <bool SCREEN, bool SAVE, int MINW>
__global__ void __launch_bounds__(256, MINW) mvs_particles_fast(ParticleArgs args) {
    extern __shared__ double lds_ends[];
    if (SAVE) {
        fast_tile<SCREEN, SAVE, -1>(args, lds_ends);
        return;
    }
    switch (args.sys[args.tiles[blockIdx.x].x].nbig - 1) {
        case 0: fast_tile<SCREEN, SAVE, 0>(args, lds_ends); break;
        case 1: fast_tile<SCREEN, SAVE, 1>(args, lds_ends); break;
        case 2: fast_tile<SCREEN, SAVE, 2>(args, lds_ends); break;
        case 3: fast_tile<SCREEN, SAVE, 3>(args, lds_ends); break;
        case 4: fast_tile<SCREEN, SAVE, 4>(args, lds_ends); break;
        case 5: fast_tile<SCREEN, SAVE, 5>(args, lds_ends); break;
        case 6: fast_tile<SCREEN, SAVE, 6>(args, lds_ends); break;
        case 7: fast_tile<SCREEN, SAVE, 7>(args, lds_ends); break;
        case 8: fast_tile<SCREEN, SAVE, 8>(args, lds_ends); break;
        case 9: fast_tile<SCREEN, SAVE, 9>(args, lds_ends); break;
        case 10: fast_tile<SCREEN, SAVE, 10>(args, lds_ends); break;
        case 11: fast_tile<SCREEN, SAVE, 11>(args, lds_ends); break;
        case 12: fast_tile<SCREEN, SAVE, 12>(args, lds_ends); break;
        default: fast_tile<SCREEN, SAVE, -1>(args, lds_ends); break;
    }
}

// Slow-set members: each runs with the whole drift chain.  The launch for
// block m (park.chunk) waits for the planet tables of its steps and for block
// m-1's fast kernel, not for block m's: it runs beside that fast kernel.  Its
// members are the kept ones and those parked by fast kernels up to block m-1
// (park.snap, or the kept count for m = 0); a member parked in block m-1 at
// step s first finishes that block (steps s.. of ptab).  The call ends with a
// launch of no steps of its own that finishes the last block's parks.
// Lanes of one wave may belong to different systems, so system and table reads
// here are per lane.
// UNIFORM: every entry belongs to system 0, so the wave walks the steps in
// lockstep from the earliest start among its lanes and reads each step's table
// with wave-uniform (scalar) loads; lanes join at their own start step.
//
// These few latency-bound waves share SIMDs with the fast kernel's; raised
// issue priority keeps them from trailing the fast launches they overlap.
//
// NG (cometary model, ngflag 1 or 3): members carry mdt_mvs's angf
// (mercury6_2.for:4054, 4059-4061, 4090-4095) in st.angf: mfo_ngf at a reset
// and after every drift, added in both kicks.  Particles with cometary or
// Yarkovsky parameters are permanent members (slow bit 2), so the fast kernel
// never sees them.
template <bool SCREEN, bool SAVE, bool UNIFORM, bool NG = false>
__global__ void __launch_bounds__(64) mvs_particles_general(ParticleArgs args) {
    __builtin_amdgcn_s_setprio(3);
    const ParkList& P = args.park;
    const int par = *P.par;
    const Park* list = P.buf + (long long)par * P.cap;
    Park* next = P.buf + (long long)(1 - par) * P.cap;
    const int count = min(P.chunk == 0 ? *P.pers : (P.live ? P.count2[par] : *P.snap), P.cap);
    const long long cap = args.st.cap;
    const int per_wave = P.per_wave;
    const int lane = threadIdx.x;          // one 64-lane wave per workgroup
    const int prev = P.chunk - 1;          // block whose parks this launch first catches up
    for (int base = blockIdx.x * per_wave; base < count; base += gridDim.x * per_wave) {
        const int idx = base + lane;
        bool act = lane < per_wave && idx < count;
        Park pk{0, 0, 0, 0};
        int a_from = args.psteps, b_from = args.steps;   // catch-up and own-block first steps (none)
        if (act) {
            pk = list[idx];
            if (pk.chunk < 0 || pk.chunk < prev) {        // kept, or parked before block m-1
                b_from = 0;
            } else if (pk.chunk == prev) {                // parked by block m-1's fast kernel
                a_from = pk.step;
                b_from = 0;
            }
            act = a_from < args.psteps || b_from < args.steps;
        }
        if (!act) { a_from = args.psteps; b_from = args.steps; }
        const Sys d = args.sys[UNIFORM ? 0 : pk.sys];
        const int g = pk.j;
        const int jb = g - d.base;
        const int np = d.nbig - 1;
        Orbit o{0, 0, 0, 0, 0, 0};
        double ax = 0, ay = 0, az = 0;
        double mj = 0.0, rce_j = 0.0, rphys_j = 0.0;
        double pg[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double ang[3] = {0.0, 0.0, 0.0};
        if (act) {
            o = Orbit{args.st.x[g], args.st.x[cap + g], args.st.x[2 * cap + g],
                      args.st.v[g], args.st.v[cap + g], args.st.v[2 * cap + g]};
            if (NG) load_ngf(args.st, g, pg);
            if (args.reset && pk.chunk < 0) {
                // a kept member starts a call that recomputes accelerations
                particle_accel(d, args.tab + d.start, np, o.x, o.y, o.z, ax, ay, az);
                if (!isfinite(ax) || !isfinite(ay) || !isfinite(az)) atomicOr(args.st.fault, FAULT_FORCE);
                if (NG)
                    ngf_body(pg[0], pg[1], pg[2], pg[4], o.x, o.y, o.z, o.vx, o.vy, o.vz, ang[0], ang[1], ang[2],
                             args.st.fault);
            } else {
                ax = args.st.a[g]; ay = args.st.a[cap + g]; az = args.st.a[2 * cap + g];
                if (NG) { ang[0] = args.st.angf[g]; ang[1] = args.st.angf[cap + g]; ang[2] = args.st.angf[2 * cap + g]; }
            }
            if (SCREEN || SAVE) { mj = args.st.m[g]; rce_j = args.st.rce[g]; rphys_j = args.st.rphys[g]; }
        }
        // the fast kernel's gate (drift_fast): the drift it would have made
        const bool in_range = d.mu >= 0x1p-100 && d.mu <= 0x1p100 && fabs(d.h) >= 0x1p-100 && fabs(d.h) <= 0x1p100;
        bool left_kepmd = false;
        // steps [from, to) of one table block; step numbers from `first`
        auto run = [&](const double* blk, int from, int to, int first) {
            const double* tab = blk + d.tab;
            int s0 = from;
            if (UNIFORM) {
                for (int off = 32; off > 0; off >>= 1) s0 = min(s0, __shfl_xor(s0, off));
                s0 = __builtin_amdgcn_readfirstlane(s0);
            }
            for (int s = s0; s < to; s++) {
                if (s < from) continue;
                const double* T = tab + (long long)s * d.ts;
                double xs[6] = {o.x, o.y, o.z, o.vx, o.vy, o.vz};
                if (NG) {
                    o.vx = o.vx + d.hby2 * ((ang[0] + 0.0) + ax);
                    o.vy = o.vy + d.hby2 * ((ang[1] + 0.0) + ay);
                    o.vz = o.vz + d.hby2 * ((ang[2] + 0.0) + az);
                } else {
                    o.vx = o.vx + d.hby2 * ax;
                    o.vy = o.vy + d.hby2 * ay;
                    o.vz = o.vz + d.hby2 * az;
                }
                Orbit k0 = o;
                const double rr = o.x * o.x + o.y * o.y + o.z * o.z;
                bool kep = false;
                drift_one(d.mu, o, d.h, &kep);
                if (!(kep && in_range && rr >= 0x1p-200 && rr <= 0x1p200)) left_kepmd = true;
                if (NG)
                    ngf_body(pg[0], pg[1], pg[2], pg[4], o.x, o.y, o.z, o.vx, o.vy, o.vz, ang[0], ang[1], ang[2],
                             args.st.fault);
                if (SAVE) save_step(args.st, g, xs, k0, o);
                // one system: the table is the same for every lane, read with scalar loads
                if constexpr (UNIFORM)
                    finish_step<SCREEN, NG>(d, const_tab(T), np, jb, mj, rce_j, rphys_j, s + first, 0, args.es,
                                            args.st.fault, xs, k0, o, ax, ay, az, ang);
                else
                    finish_step<SCREEN, NG>(d, T, np, jb, mj, rce_j, rphys_j, s + first, pk.sys, args.es,
                                            args.st.fault, xs, k0, o, ax, ay, az, ang);
            }
        };
        if (args.psteps > 0) run(args.ptab, a_from, args.psteps, args.step0 - args.psteps);
        run(args.tab, b_from, args.steps, args.step0);
        if (act) {
            args.st.x[g] = o.x; args.st.x[cap + g] = o.y; args.st.x[2 * cap + g] = o.z;
            args.st.v[g] = o.vx; args.st.v[cap + g] = o.vy; args.st.v[2 * cap + g] = o.vz;
            args.st.a[g] = ax; args.st.a[cap + g] = ay; args.st.a[2 * cap + g] = az;
            if (NG) { args.st.angf[g] = ang[0]; args.st.angf[cap + g] = ang[1]; args.st.angf[2 * cap + g] = ang[2]; }
            if (args.work)
                atomicAdd(args.work + 1, (unsigned long long)(args.psteps - a_from + args.steps - b_from));
            // members parked in this call stay; kept members stay while they leave the kepmd branch;
            // permanent members always stay
            const bool kept = pk.chunk < 0;
            const unsigned char bits = P.slow[g];
            if (P.last) {
                if (!kept || left_kepmd || (bits & 6)) {
                    int slot = atomicAdd(P.count2 + (1 - par), 1);
                    if (slot < P.cap) next[slot] = Park{g, 0, pk.sys, -1};
                    P.slow[g] = (unsigned char)(1 | (bits & 4));
                } else {
                    P.leave[atomicAdd(P.nleave, 1)] = g;
                }
            } else if (kept && left_kepmd) {
                P.slow[g] = (unsigned char)(bits | 2);
            }
        }
    }
}

// Host-driven event screen over the step-start copies of one single-step
// launch (the reference's events call, mercury_cuda.cu:470-493): pairs with
// the Big bodies of step 0's table and central impacts, every body j >= 1.
struct SavedScreenArgs {
    const Sys* sys;
    State st;
    const double* tab;
    EventSink es;
    double time, h, rcen;
    int pairs;          // 0: central impacts only
    const double* va;   // pair-screen velocities at the start and end of the step
    const double* vb;   // (MVS: after the first kick / after the drift; others: ov, v)
    const double* vc;   // central check: velocity at the end (HYBRID: the physical one)
};

__global__ void __launch_bounds__(256) screen_saved(SavedScreenArgs args) {
    const Sys d = args.sys[0];
    const int jb = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (jb >= d.n) return;
    const long long cap = args.st.cap;
    const int g = d.base + jb;
    const int np = d.nbig - 1;
    const double* T = args.tab + d.tab;
    const double mj = args.st.m[g];
    double x1 = args.st.x[g], y1 = args.st.x[cap + g], z1 = args.st.x[2 * cap + g];
    double xs = args.st.ox[g], ys = args.st.ox[cap + g], zs = args.st.ox[2 * cap + g];
    if (args.pairs) {
        Ends e{xs, ys, zs, args.va[g], args.va[cap + g], args.va[2 * cap + g],
               x1, y1, z1, args.vb[g], args.vb[cap + g], args.vb[2 * cap + g]};
        screen_body(T, np, jb, d.nbig, mj, e, args.st.rce[g], args.st.rphys[g], args.h, args.time, 0, 0,
                    args.es);
    }
    central_check(d.mu, mj, args.rcen, args.h, args.time, xs, ys, zs, args.st.ov[g], args.st.ov[cap + g],
                  args.st.ov[2 * cap + g], x1, y1, z1, args.vc[g], args.vc[cap + g],
                  args.vc[2 * cap + g], jb, 0, 0, args.es);
}

// Details of the fast kernel's deferred screening work (one lane per request).
struct RequestArgs {
    const Sys* sys;
    State st;
    const double* tab;
    EventSink es;
    ReqSink req;
    int step0;
};

__global__ void __launch_bounds__(256) screen_requests(RequestArgs args) {
    const int count = min(*args.req.count, args.req.cap);
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < count; idx += gridDim.x * blockDim.x) {
        const ScreenReq r = args.req.buf[idx];
        const Sys d = args.sys[r.sys];
        const int np = d.nbig - 1;
        const double* T = args.tab + d.tab + (long long)r.step * d.ts;
        const double mj = args.st.m[r.g];
        if (r.kind == 0) {
            const double* F = T + kHdr;
            const double* B = F + kF * np + kScr * r.p;
            Ends e{r.w[0], r.w[1], r.w[2], r.w[3], r.w[4], r.w[5], r.w[6], r.w[7], r.w[8], r.w[9], r.w[10], r.w[11]};
            screen_pair(F + kF * r.p, B, r.p + 1, r.jb, r.jb < d.nbig, F[kF * r.p + 3], mj, e,
                        args.st.rce[r.g], args.st.rphys[r.g], d.h, T[6], r.step + args.step0, r.sys, args.es);
        } else {
            double rr0 = r.w[0] * r.w[0] + r.w[1] * r.w[1] + r.w[2] * r.w[2];
            double rr1 = r.w[6] * r.w[6] + r.w[7] * r.w[7] + r.w[8] * r.w[8];
            double rv0 = r.w[3] * r.w[0] + r.w[4] * r.w[1] + r.w[5] * r.w[2];
            central_event(d.mu, mj, d.rcen, d.h, T[6], r.w[0], r.w[1], r.w[2], r.w[3], r.w[4], r.w[5],
                          rr0, rr1, rv0, r.jb, r.step + args.step0, r.sys, args.es);
        }
    }
}

__global__ void zero_ints(int* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void zero_bytes(unsigned char* p, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void copy_int(const int* src, int* dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

// An empty kernel: a point in a stream's order that an event can mark.
__global__ void marker() {}

// Current slow-set size after a fast launch (the general launch of the same
// chunk takes entries up to it; later fast launches keep appending).
__global__ void snap_count(const int* count2, const int* par, int* dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = count2[*par];
}

// call_begin_body as a launch of its own (ensembles: their planet launches
// run several workgroups at once).
__global__ void __launch_bounds__(256) call_begin(CallBeginArgs a) { call_begin_body(a); }

// End of a call, after every stream's work: the counts the host reads, written
// into pinned host memory so that no copy follows the call.
// status: [0] fault flags, [1] events, [2] slow-set members in the call,
// [3] members kept for the next call.
__global__ void call_end(const int* count2, const int* par, const int* devcount, const int* fault,
                         int* status) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int q = *par;
        status[0] = *fault;
        status[1] = *devcount;
        status[2] = count2[q];
        status[3] = count2[1 - q];
        __threadfence_system();
    }
}

}  // namespace mhip
