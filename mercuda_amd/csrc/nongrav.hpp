// Non-gravitational accelerations for gfx950, FP64, one body per lane:
//   ngf_body  mfo_ngf (mercury6_2.for:5106-5152): the cometary A1 (radial),
//             A2 (transverse in the orbit plane) and A3 (normal) terms scaled by
//             Marsden, Sekanina and Yeomans' g(r) (AJ 78, 211, 1973), switched
//             off beyond r^2 = 88 AU^2 unless some |A| exceeds 1e-7; and the
//             Yarkovsky term yar / r^2 along the transverse direction;
//   pr_body   mfo_pr (mercury6_2.for:5307-5357): radiation pressure and
//             Poynting-Robertson drag with a 30% solar-wind share (Burns, Lamy
//             and Soter, Icarus 40, 1, 1979) on massless bodies, beta = ngf(4).
// Both keep the Fortran's operation order; ngf holds A1, A2, A3, beta, yar.
//
// g(r) = 0.111262 q^-2.15 (1 + q^5.093)^-4.6142, q = r / 2.808 AU, needs x^y.
// pow_cr returns the correctly rounded x^y: log and exp carried in
// double-double (about 2^-100 relative), then rounded once.  The CPU path
// calls libm's pow, which is within 0.52 ulp and so differs from the correctly
// rounded value in the last bit for about one argument in a thousand; such a
// one-ulp difference in g moves a cometary acceleration by ~1e-16 of itself,
// ~1e-20 of the gravity it is added to, and the sum almost always rounds the
// same.
#pragma once
#include <hip/hip_runtime.h>

namespace mhip {

enum { FAULT_NGF = 8 };   // undefined transverse or normal direction (mfo_ngf's fail)

constexpr double kK2 = 2.959122082855911e-4;                                   // mercury.inc:41
constexpr double kCLight = 299792458.0 * 86400.0 / (1.495978707e13 * .01);     // AU/day, mfo_pr

// double-double: value hi + lo with |lo| <= ulp(hi) / 2
struct DD {
    double hi, lo;
};

__host__ __device__ __forceinline__ DD dd_fast(double a, double b) {   // |a| >= |b| or a == 0
    double s = a + b;
    return DD{s, b - (s - a)};
}
__host__ __device__ __forceinline__ DD dd_sum(double a, double b) {    // Knuth's two-sum
    double s = a + b;
    double bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD s = dd_sum(a.hi, b.hi);
    DD t = dd_sum(a.lo, b.lo);
    s = dd_fast(s.hi, s.lo + t.hi);
    return dd_fast(s.hi, s.lo + t.lo);
}
__host__ __device__ __forceinline__ DD dd_mul(DD a, DD b) {
    double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e = e + (a.hi * b.lo + a.lo * b.hi);
    return dd_fast(p, e);
}
__host__ __device__ __forceinline__ DD dd_mul_d(DD a, double b) {
    double p = a.hi * b;
    double e = fma(a.hi, b, -p);
    e = e + a.lo * b;
    return dd_fast(p, e);
}
__host__ __device__ __forceinline__ DD dd_div_d(DD a, double b) {
    double q1 = a.hi / b;
    double p = q1 * b;
    double e = fma(q1, b, -p);              // q1 b = p + e exactly
    double r = ((a.hi - p) - e) + a.lo;     // a - q1 b
    return dd_fast(q1, r / b);
}

constexpr DD kLn2{0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};

// e^r - 1 for |r| <= 0.36: Taylor to degree 9 of r / 2^10 (remainder below
// 2^-120 relative), then ten squarings (1 + u)^2 - 1 = 2u + u^2.
__host__ __device__ inline DD dd_expm1_small(DD r) {
    const DD s{ldexp(r.hi, -10), ldexp(r.lo, -10)};
    DD p{1.0, 0.0};
    for (int k = 9; k >= 2; k--) p = dd_add(DD{1.0, 0.0}, dd_div_d(dd_mul(s, p), (double)k));
    DD u = dd_mul(s, p);
    for (int i = 0; i < 10; i++) u = dd_add(DD{2.0 * u.hi, 2.0 * u.lo}, dd_mul(u, u));
    return u;
}

// log x for normal positive x: x = f 2^k with f in [sqrt(1/2), sqrt(2));
// y0 = log f to a few ulp, then one Newton step in double-double:
// log f = y0 + log1p(f e^-y0 - 1), the argument t ~ 2^-52 (t^3 dropped).
__host__ __device__ inline DD dd_log(double x) {
    int k;
    double f = frexp(x, &k);
    if (f < 0.70710678118654752) {
        f = f * 2.0;
        k = k - 1;
    }
    const double y0 = log(f);
    const DD u = dd_expm1_small(DD{-y0, 0.0});
    const DD t = dd_add(DD{f - 1.0, 0.0}, dd_mul_d(u, f));   // f - 1 is exact (Sterbenz)
    const DD l = dd_add(DD{y0, 0.0}, dd_add(t, DD{-0.5 * t.hi * t.hi, 0.0}));
    return dd_add(dd_mul_d(kLn2, (double)k), l);
}

// x^y correctly rounded for x > 0 with x and the result normal
// (|y log x| < 700); outside that range the library pow.
__host__ __device__ inline double pow_cr(double x, double y) {
    if (!(x >= 0x1p-1000) || !(x <= 0x1p1000)) return pow(x, y);
    const DD z = dd_mul_d(dd_log(x), y);
    if (!(fabs(z.hi) < 700.0)) return pow(x, y);
    const double kd = rint(z.hi * 1.4426950408889634);
    const DD r = dd_add(z, dd_mul_d(DD{-kLn2.hi, -kLn2.lo}, kd));
    const DD e = dd_add(DD{1.0, 0.0}, dd_expm1_small(r));
    return ldexp(e.hi, (int)kd);
}

// Marsden's g(r), r in AU (mercury6_2.for:5130-5131)
__host__ __device__ inline double marsden_g(double r) {
    const double q = r * .3561253561253561;
    return .111262 * pow_cr(q, -2.15) * pow_cr(1.0 + pow_cr(q, 5.093), -4.6142);
}

// mfo_ngf for one body j >= 1 with parameters a1, a2, a3, yar: its
// acceleration (zero when all four are zero).  An undefined direction sets
// FAULT_NGF (the Fortran stops there).
__device__ inline void ngf_body(double a1, double a2, double a3, double yar, double x, double y, double z,
                                double vx, double vy, double vz, double& ax, double& ay, double& az, int* fault) {
    ax = 0.0; ay = 0.0; az = 0.0;
    if (a1 == 0.0 && a2 == 0.0 && a3 == 0.0 && yar == 0.0) return;
    const double r2 = x * x + y * y + z * z;
    const double r = sqrt(r2);
    const double rv = (x * vx + y * vy + z * vz) / r2;
    const double tx = vx - rv * x;
    const double ty = vy - rv * y;
    const double tz = vz - rv * z;
    const double tn = sqrt(tx * tx + ty * ty + tz * tz);
    if ((a1 != 0.0 || a2 != 0.0 || a3 != 0.0) &&
        (r2 < 88.0 || fabs(a1) > 1e-7 || fabs(a2) > 1e-7 || fabs(a3) > 1e-7)) {
        const double g = marsden_g(r);
        const double f1 = a1 * g / r;
        ax = f1 * x; ay = f1 * y; az = f1 * z;
        if (a2 != 0.0) {
            if (tn == 0.0) atomicOr(fault, FAULT_NGF);
            const double f = a2 * g / tn;
            ax = ax + f * tx; ay = ay + f * ty; az = az + f * tz;
        }
        if (a3 != 0.0) {
            const double nx = y * vz - z * vy;
            const double ny = z * vx - x * vz;
            const double nz = x * vy - y * vx;
            const double nn = sqrt(nx * nx + ny * ny + nz * nz);
            if (nn == 0.0) atomicOr(fault, FAULT_NGF);
            const double f = a3 * g / nn;
            ax = ax + f * nx; ay = ay + f * ny; az = az + f * nz;
        }
    }
    if (yar != 0.0) {
        if (tn == 0.0) atomicOr(fault, FAULT_NGF);
        const double f = yar / (r2 * tn);
        ax = ax + f * tx; ay = ay + f * ty; az = az + f * tz;
    }
}

// mfo_pr for one body of mass mj (G m) with beta: zero for a massive body.
__device__ inline void pr_body(double mj, double beta, double x, double y, double z, double vx, double vy, double vz,
                               double& ax, double& ay, double& az) {
    ax = 0.0; ay = 0.0; az = 0.0;
    if (mj != 0.0 || beta == 0.0) return;
    const double sw = .3;
    const double r = sqrt(x * x + y * y + z * z);
    const double vr = (x * vx + y * vy + z * vz) / r;
    const double f = kK2 * beta / (r * r);
    const double w = 1.0 - (1.0 + sw) * vr / kCLight;
    ax = f * (w * x / r - (1.0 + sw) * vx / kCLight);
    ay = f * (w * y / r - (1.0 + sw) * vy / kCLight);
    az = f * (w * z / r - (1.0 + sw) * vz / kCLight);
}

// The non-gravitational part of mfo_all (mercury6_2.for:4560-4580) added to
// a body's acceleration: mfo_ngf's when flag is 1 or 3, then mfo_pr's when 2
// or 3, each added as the Fortran adds its acor array (zeros included).
// P: the body's five parameters A1, A2, A3, beta, yar.
__device__ inline void add_nongrav(int flag, const double P[5], double mj, double x, double y, double z, double vx,
                                   double vy, double vz, double& ax, double& ay, double& az, int* fault) {
    double cx, cy, cz;
    if (flag == 1 || flag == 3) {
        ngf_body(P[0], P[1], P[2], P[4], x, y, z, vx, vy, vz, cx, cy, cz, fault);
        ax = ax + cx; ay = ay + cy; az = az + cz;
    }
    if (flag == 2 || flag == 3) {
        pr_body(mj, P[3], x, y, z, vx, vy, vz, cx, cy, cz);
        ax = ax + cx; ay = ay + cy; az = az + cz;
    }
}

}  // namespace mhip
