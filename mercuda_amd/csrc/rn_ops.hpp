// Correctly rounded FP64 divide and square root for gfx950 without the range
// scaling of hipcc's general expansion.
//
// hipcc expands a / b as v_div_scale (twice), v_rcp_f64, two Newton steps,
// the product q = a*r, its residual a - b*q, v_div_fmas and v_div_fixup: 11
// VALU ops.  When |a|, |b| and |a/b| lie far inside the normal range
// (2^-900 < |a|, 2^-1000 < |b| < 2^1000, |a/b| in [2^-1000, 2^760]) the two
// v_div_scale return their operands unchanged, v_div_fmas is a plain fma and
// v_div_fixup passes its value through, so div_rn() -- the same arithmetic
// without those four ops -- returns the same bits in 8 ops.  The one
// difference inside that range is the sign of a zero quotient from a -0
// numerator (+0 here), which no caller below can observe: each such quotient
// is added to a nonzero term before it is used.
//
// sqrt(x) is v_rsq_f64 and a Goldschmidt refinement (10 ops), wrapped in a
// scale of inputs below 2^-767 and a pass-through of ±0 and +inf (9 more
// ops); sqrt_rn() is the core, exact for x in [2^-767, 2^1023].
//
// divc_rn(a, b, y) divides by a constant b with y = RN(1/b) precomputed:
// q = a*y is within 1 ulp of a/b, and Markstein's correction
// q + (a - b*q)*y with fused operations rounds to RN(a/b) (Markstein, IBM J.
// Res. Dev. 34 (1990); Muller et al., Handbook of Floating-Point Arithmetic,
// ch. 4), 3 ops in place of 11.
//
// The device self-check mercury_hip_check_arith compares all three with the
// compiler's expansions over random operands of those ranges.
#pragma once
#include <hip/hip_runtime.h>

namespace mhip {

__device__ __forceinline__ double div_rn(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    double q = a * r;
    double t = __builtin_fma(-b, q, a);
    return __builtin_fma(t, r, q);
}

// div_rn split at its divisor: rcp_refine(b) is the Newton-refined reciprocal
// that div_rn computes from b alone, div_rn_r(a, b, r) the rest, so several
// quotients by one divisor share the refinement with div_rn's bits.
__device__ __forceinline__ double rcp_refine(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double div_rn_r(double a, double b, double r) {
    double q = a * r;
    double t = __builtin_fma(-b, q, a);
    return __builtin_fma(t, r, q);
}

__device__ __forceinline__ double rcp_rn(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    double t = __builtin_fma(-b, r, 1.0);
    return __builtin_fma(t, r, r);
}

__device__ __forceinline__ double sqrt_rn(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

__device__ __forceinline__ double divc_rn(double a, double b, double y) {
    double q = a * y;
    double t = __builtin_fma(-b, q, a);
    return __builtin_fma(t, y, q);
}

// Operation sets for code shared by the fast kernel (RangeOps, operands
// bounded as above) and every other path (IeeeOps, the compiler's IEEE
// expansions).  div_c(a, b, y) divides by the constant b with y = RN(1/b).
// sq_ok(s2): a squared distance whose 1/sqrt stays exact; outside it (RangeOps)
// s^-3 has overflowed or s^2 itself has, as it would in the IEEE form too.
struct IeeeOps {
    static __device__ __forceinline__ bool sq_ok(double s2) { return s2 > 0.0; }
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double rcp(double b) { return 1.0 / b; }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double div_c(double a, double b, double) { return a / b; }
};

struct RangeOps {
    static __device__ __forceinline__ bool sq_ok(double s2) { return s2 > 0x1p-767 && s2 < 0x1p1000; }
    static __device__ __forceinline__ double div(double a, double b) { return div_rn(a, b); }
    static __device__ __forceinline__ double rcp(double b) { return rcp_rn(b); }
    static __device__ __forceinline__ double sqrt(double x) { return sqrt_rn(x); }
    static __device__ __forceinline__ double div_c(double a, double b, double y) { return divc_rn(a, b, y); }
};

// Self-check: each lane draws operands from a counter-based generator and
// compares every operation above with the compiler's expansion, bit for bit.
// bad[0..3] count mismatches of div_rn, rcp_rn, sqrt_rn and divc_rn.
//   div_rn:  |a| in [2^-450, 2^450], |b| in [2^-300, 2^300], random signs
//   rcp_rn:  |b| in [2^-700, 2^700]
//   sqrt_rn: x in [2^-767, 2^1000]
//   divc_rn: |a| in [2^-900, 2^900] over 3, 2*pi and 39916800
// A quarter of the draws take edge significands (see draw).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// A random exponent in [emin, emax]; the significand random, or for one draw
// in four one of the edge patterns where a Newton reciprocal can round to the
// wrong neighbour: near all ones, near a power of two, few low bits.
__device__ __forceinline__ double draw(unsigned long long r, int emin, int emax, bool sign) {
    unsigned long long e = (unsigned long long)(emin + (int)((r >> 52) % (unsigned long long)(emax - emin + 1)) + 1023);
    const unsigned long long mask = (1ull << 52) - 1;
    unsigned long long f = r & mask;
    switch ((r >> 62) == 3 ? (int)((r >> 32) & 7) : -1) {
        case 0: f = mask; break;                               // 1.111...1
        case 1: f = mask - (r & 0xffff); break;                // just below 2
        case 2: f = 0; break;                                  // a power of two
        case 3: f = r & 0xffff; break;                         // just above a power of two
        case 4: f = (r & 0xff) << 44; break;                   // few leading bits
        case 5: f = mask ^ ((r & 0xff) << 44); break;          // ones with a few holes
        default: break;
    }
    double v = __longlong_as_double((long long)((e << 52) | f));
    return sign ? -v : v;
}

__global__ void __launch_bounds__(256) check_arith(long long n, unsigned long long seed,
                                                   unsigned long long* bad) {
    unsigned long long cnt[4] = {0, 0, 0, 0};
    const double cb[3] = {3.0, 2.0 * 3.141592653589793, 39916800.0};
    const double cy[3] = {1.0 / 3.0, 1.0 / (2.0 * 3.141592653589793), 1.0 / 39916800.0};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        unsigned long long r0 = mix64(seed ^ (unsigned long long)(4 * i));
        unsigned long long r1 = mix64(seed ^ (unsigned long long)(4 * i + 1));
        unsigned long long r2 = mix64(seed ^ (unsigned long long)(4 * i + 2));
        unsigned long long r3 = mix64(seed ^ (unsigned long long)(4 * i + 3));
        double a = draw(r0, -450, 450, r2 & 1);
        double b = draw(r1, -300, 300, r2 & 2);
        if (__double_as_longlong(div_rn(a, b)) != __double_as_longlong(a / b)) cnt[0]++;
        double c = draw(r2, -700, 700, r3 & 1);
        if (__double_as_longlong(rcp_rn(c)) != __double_as_longlong(1.0 / c)) cnt[1]++;
        double x = draw(r3, -767, 1000, false);
        if (__double_as_longlong(sqrt_rn(x)) != __double_as_longlong(sqrt(x))) cnt[2]++;
        double w = draw(r0 ^ r3, -900, 900, r3 & 2);
        int k = (int)((r1 & 0xffff) % 3);
        if (__double_as_longlong(divc_rn(w, cb[k], cy[k])) != __double_as_longlong(w / cb[k])) cnt[3]++;
    }
    for (int q = 0; q < 4; q++)
        if (cnt[q]) atomicAdd(bad + q, cnt[q]);
}

}  // namespace mhip
