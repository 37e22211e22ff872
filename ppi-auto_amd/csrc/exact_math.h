// exact_math.h — float-in / float-out exp, log, sin and cos on gfx950, rounded
// exactly as (float) f((double) x).
//
// The reference's sequential path evaluates these operators in double and rounds
// the result to float (functions.h, interpreter_core:103-117,140-142: the float
// argument promotes to the double routine).  The ROCm device library's double
// routines reach double accuracy with long polynomials and double-double
// corrections (40-100 VALU instructions each on gfx950).  A float result needs far
// less: it only has to be rounded correctly.
//
// Method (Tang's table-driven reductions, with Ziv's rounding test):
//   1. reduce x with a small table so that a short polynomial is accurate, and
//      compute y ~ f(x) in double with a relative error below 2^-50 (the bounds
//      are worked out per function below);
//   2. (float) y is the correctly rounded float of f(x) unless f(x) lies within
//      that error of a rounding boundary (the midpoint between two floats).  The
//      boundary sits where bits 0-28 of y's significand are 0x10000000; y is
//      accepted when those bits are more than kGuard = 64 units away from it
//      (64 double ulps > 2^-47 relative, 8x the worst error);
//   3. otherwise — about one evaluation in 4 million, and every input outside
//      the reduced domain (NaN, inf, huge arguments, denormal results) — the
//      device library's double routine is called and rounded, as before.
// When the test passes, the device library's result (within 1 double ulp of
// f(x)) rounds to the same float, so every result is identical to rounding the
// device library's double routine; the GPU test compares the two on all 2^32
// float inputs of each function.
//
// Tables and polynomial coefficients: exact_tables.h (generated by
// gen_exact_tables.py with mpmath).  Every step below is written with explicit
// fma() and no contraction, so the code computes exactly what is written.
#ifndef PPI_EXACT_MATH_H
#define PPI_EXACT_MATH_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "exact_tables.h"

namespace ppi {
namespace dev {

// The device library's double routines, rounded to float: the fallback of the
// fast paths, called out of line under a branch that is rarely taken.
__device__ __noinline__ float exp_of_double(float x) { return (float)exp((double)x); }
__device__ __noinline__ float log_of_double(float x) { return (float)log((double)x); }
__device__ __noinline__ float sin_of_double(float x) { return (float)sin((double)x); }
__device__ __noinline__ float cos_of_double(float x) { return (float)cos((double)x); }
__device__ __noinline__ float tan_of_double(float x) { return (float)tan((double)x); }
__device__ __noinline__ float sqrt_float(float x) { return sqrtf(x); }

constexpr uint32_t kGuard = 64;  // double ulps kept between y and a float rounding boundary

// The tables live in LDS (4 KB per workgroup): a table read is a ds_read with
// LDS latency instead of a vector-memory load.  The polynomial coefficients are
// scalar constants; a gfx9 VALU instruction takes at most one scalar operand,
// so each polynomial's first step is a multiply and an add rather than an fma
// with two constant operands (which would need a v_mov pair).  Every kernel that evaluates the
// exact routines stages them once at entry — stage_exact_tables() by all
// threads of the workgroup, then a barrier — before any thread reads them.
// 64-byte alignment puts the tables at LDS address 0, ahead of the kernels' feature tiles: every
// table read then takes its base from the instruction's offset field (the sine pair's ds_read2_b64
// has only 8-bit offsets), one VALU add fewer per sin, cos and log row.
__shared__ __attribute__((aligned(64))) ExactTables lds_exact_tables;

__device__ __forceinline__ void stage_exact_tables() {
    constexpr int kWords = (int)(sizeof(ExactTables) / sizeof(double));
    const double* src = reinterpret_cast<const double*>(&kExactTables);
    double* dst = reinterpret_cast<double*>(&lds_exact_tables);
    for (int i = (int)threadIdx.x; i < kWords; i += (int)blockDim.x) dst[i] = src[i];
}

// Ziv's test for rounding a double y (with |y| in the float normal range, or 0)
// to float: true when bits 0-28 of y's significand (lo) are more than kGuard away
// from the midpoint pattern 0x10000000, i.e. lo outside [2^28 - G, 2^28 + G].
// With w = (lo + 2^28 + G) mod 2^29 that is w > 2G (lo below the band wraps to
// w >= 2^28 + G), and 8w = (8 y.lo + 2^31 + 8G) mod 2^32 exactly: one shift-add
// and one compare per row instead of a mask, an add and a compare.
__device__ __forceinline__ bool rounds_safely(double y) {
    const uint32_t w8 = (uint32_t)__double2loint(y) * 8u + (0x80000000u + 8u * kGuard);
    return w8 > 16u * kGuard;
}

// exp.  x = k ln2/128 + r with |r| <= ln2/256 (+ rounding); exp(x) = 2^(k>>7) T[k&127] (1 + p(r)).
// Error: r is exact to 2^-62 absolute (the first fma is exact, the second rounds once, the tail of
// ln2/128 beyond kExpL1 + kExpL2 times |k| < 2^15 is below 2^-99); truncating exp(r) - 1 after r^4
// leaves r^5/120 < 2^-49.6; T[j] and the last fma each add 2^-53.  Total < 2^-49.3 relative.
// Domain: -87 < x < 128 (float results in the normal range or overflowing to inf); x >= 128 gives
// +inf and x <= -104 gives +0 directly (exp(-104) < 2^-150, half the smallest denormal).
__device__ __forceinline__ float exact_exp_fast(float x, bool& ok) {
    const double xd = (double)x;
    const double k = __builtin_rint(xd * kExpInvL);
    const int ki = (int)k;
    double r = __builtin_fma(-k, kExpL1, xd);
    r = __builtin_fma(-k, kExpL2, r);
    const double t = lds_exact_tables.exp2j[ki & 127];
    double q = r * kE4 + kE3;  // mul + add: one scalar operand per instruction
    q = __builtin_fma(r, q, kE2);
    const double p = __builtin_fma(r * r, q, r);
    const double y = __builtin_ldexp(__builtin_fma(t, p, t), ki >> 7);
    float res = (float)y;
    if (x >= 128.0f) res = __builtin_inff();
    if (x <= -104.0f) res = 0.0f;
    // Bitwise rather than short-circuit operators: the same value, without the exec-mask
    // branches the compiler builds around a short-circuited rounds_safely.
    ok = ((x > -87.0f) & rounds_safely(y)) | (x >= 128.0f) | (x <= -104.0f);
    return res;
}

// log.  x = 2^k z with z in [0.69921875, 1.3984375); the bucket i of z gives invc ~ 1/c and
// logc = -log(invc); log(x) = k ln2 + logc + log1p(r), r = z invc - 1 (one rounding: the product
// of a 24-bit z and a double is exact inside the fma).  |r| <= 2^-7; truncating log1p after r^7
// leaves r^7/8 < 2^-52 relative.  The two buckets around z = 1 use invc = 1, so log(x) ~ r there
// carries no cancellation; elsewhere |k ln2 + logc| and |log1p r| are at most 2x |log x|, so with
// the four roundings (table, two fmas, sum) the error stays below 2^-50 relative.  Domain:
// positive normal x.
__device__ __forceinline__ float exact_log_fast(float x, bool& ok) {
    const uint32_t ix = __float_as_uint(x);
    const uint32_t tmp = ix - 0x3f330000u;
    const int k = (int)tmp >> 23;
    const uint32_t i = (tmp >> 16) & 127u;
    const float z = __uint_as_float(ix - (tmp & 0xff800000u));
    const double r = __builtin_fma((double)z, lds_exact_tables.log[i][0], -1.0);
    double q = r * kL7 + kL6;
    q = __builtin_fma(r, q, kL5);
    q = __builtin_fma(r, q, kL4);
    q = __builtin_fma(r, q, kL3);
    q = __builtin_fma(r, q, kL2);
    const double p = __builtin_fma(r * r, q, r);
    const double kd = (double)k;
    const double a = __builtin_fma(kd, kLn2Hi, lds_exact_tables.log[i][1]);
    const double b = __builtin_fma(kd, kLn2Lo, p);
    const double y = a + b;
    float res = (float)y;
    ok = (ix - 0x00800000u < 0x7f000000u) && rounds_safely(y);  // 2^-126 <= x < inf
    return res;
}

// sin and cos.  a = |x| = k pi/128 + r, |r| <= pi/256 (+ rounding); with S_j = sin(j pi/128),
// sin(a) = S_j cos r + S_{j+64} sin r (j = k mod 256) and cos(a) = sin(a + pi/2) takes j + 64
// (the table runs to j = 383, so both entries sit at fixed offsets from one index).
// r: the first fma is exact (a has at most 24 significant bits, k < 2^29.4 on the domain, and
// a - k kTrigP1 needs at most 52 bits), the second rounds once, and the tail of pi/128 beyond
// kTrigP1 + kTrigP2 times k is below 2^-82; a float below 2^24 never comes closer than 2^-26.8 to
// a multiple of pi (checked by enumerating the multiples), so where the result is ~ +-r (S_j = 0)
// r is relative-exact to 2^-53.  Elsewhere |result| >= sin(pi/256) and the terms are at most 2x
// the result.  Truncation: sin r after r^5 leaves r^7/7! < 2^-50.4 relative, cos r - 1 after r^6
// leaves r^8/8! < 2^-60.  Total below 2^-49.  Domain: |x| < 2^24.
template <bool COS>
__device__ __forceinline__ float exact_sincos_fast(float x, bool& ok) {
    const float ax = __builtin_fabsf(x);
    const double a = (double)ax;
    const double k = __builtin_rint(a * kTrigInvP);
    const int ki = (int)k;
    double r = __builtin_fma(-k, kTrigP1, a);
    r = __builtin_fma(-k, kTrigP2, r);
    const double* T = lds_exact_tables.sinpi128 + (ki & 255) + (COS ? 64 : 0);
    const double S = T[0], C = T[64];
    const double r2 = r * r;
    const double s = __builtin_fma(r2 * r, r2 * kS5 + kS3, r);
    const double cm1 = r2 * __builtin_fma(r2, r2 * kC6 + kC4, kC2);
    const double y = __builtin_fma(S, cm1, __builtin_fma(C, s, S));
    float res = (float)y;
    if (!COS) res = __uint_as_float(__float_as_uint(res) ^ (__float_as_uint(x) & 0x80000000u));  // sin is odd
    ok = ax < 16777216.0f && rounds_safely(y);
    return res;
}

// sqrt (functions.h:115-129: float(sqrt(double(x))) == sqrtf(x), the correctly rounded float
// square root).  v_sqrt_f32 is within 1 ulp; the correctly rounded result is that value or a
// neighbour, chosen by the signs of the exact residuals x - s'*s of the two neighbours (one fma
// each) — the compiler's own sqrtf expansion, without its scaling of inputs below 2^-96 (whose
// residuals could leave the normal range) and its handling of 0, inf and NaN, which stay with
// the fallback.  Domain: 2^-96 <= x < inf.
__device__ __forceinline__ float exact_sqrt_fast(float x, bool& ok) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u), sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float rm = __builtin_fmaf(-sm, s, x), rp = __builtin_fmaf(-sp, s, x);
    float res = rm <= 0.0f ? sm : s;
    res = rp > 0.0f ? sp : res;
    ok = x >= 0x1.0p-96f && x < __builtin_inff();
    return res;
}

// Function ids of the wrappers below.
enum : int { kFnExp = 0, kFnLog = 1, kFnSin = 2, kFnCos = 3, kFnSqrt = 4 };

template <int FN>
__device__ __forceinline__ float exact_fast(float x, bool& ok) {
    if constexpr (FN == kFnExp) return exact_exp_fast(x, ok);
    else if constexpr (FN == kFnLog) return exact_log_fast(x, ok);
    else if constexpr (FN == kFnSqrt) return exact_sqrt_fast(x, ok);
    else return exact_sincos_fast<FN == kFnCos>(x, ok);
}
template <int FN>
__device__ __forceinline__ float exact_fallback(float x) {
    if constexpr (FN == kFnExp) return exp_of_double(x);
    else if constexpr (FN == kFnLog) return log_of_double(x);
    else if constexpr (FN == kFnSin) return sin_of_double(x);
    else if constexpr (FN == kFnSqrt) return sqrt_float(x);
    else return cos_of_double(x);
}

// One value.
template <int FN>
__device__ __forceinline__ float exact_one(float x) {
    bool ok;
    float res = exact_fast<FN>(x, ok);
    if (!ok) res = exact_fallback<FN>(x);
    return res;
}
__device__ __forceinline__ float exact_exp(float x) { return exact_one<kFnExp>(x); }
__device__ __forceinline__ float exact_log(float x) { return exact_one<kFnLog>(x); }
__device__ __forceinline__ float exact_sin(float x) { return exact_one<kFnSin>(x); }
__device__ __forceinline__ float exact_cos(float x) { return exact_one<kFnCos>(x); }
__device__ __forceinline__ float exact_sqrt(float x) { return exact_one<kFnSqrt>(x); }

// R values of one lane (the interpreters' rows): every fast path runs first, so
// the R table loads are in flight together, and a single rarely taken branch
// holds the fallbacks.
template <int FN, int R>
__device__ __forceinline__ void exact_rows(float (&v)[R]) {
    bool ok[R];
    float res[R];
    bool all = true;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        res[r] = exact_fast<FN>(v[r], ok[r]);
        all = all && ok[r];
    }
    if (!all) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (!ok[r]) res[r] = exact_fallback<FN>(v[r]);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = res[r];
}

}  // namespace dev
}  // namespace ppi

#endif
