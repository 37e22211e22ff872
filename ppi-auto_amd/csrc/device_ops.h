// device_ops.h — operator semantics on gfx950.
//
// EXACT mode reproduces the reference's sequential path, which is the oracle:
// sequential.cc includes functions.h and interpreter_core in a C++ unit whose
// sin/cos/exp/log/pow/fmod/sinh/tan/exp10/exp2/log10/log2 calls resolve to the
// double routines, so every transcendental is evaluated in double and rounded
// to float when it is stored on the float stack.  Here exp, log, sin and cos
// are the correctly rounding float routines of exact_math.h (identical results
// to rounding the ROCm device library's double routines, at a fraction of the
// cost); the other transcendentals call the device library's double routines
// and round.  Basic arithmetic is float with no FMA contraction (the file is
// compiled with -ffp-contract=off).
//
// NATIVE mode mirrors the reference's -DNATIVE OpenCL build (functions.h:59-193,
// native_* built-ins): fast float intrinsics for exactly the operators that
// functions.h switches (SIN, COS, EXP, EXP10, EXP2, DIV, LOG, LOG10, LOG2, TAN;
// SQRT stays exact, which native_sqrt may be), not bit-compatible with the
// sequential path.  POW, GAMMA's sinh and pow, and MOD call pow, sinh and fmod
// directly in interpreter_core:68-75,127-129, which NATIVE does not touch, so
// they keep the sequential path's double routines in both modes.  The
// sequential drop-in never selects NATIVE: its native_* shims are the double
// routines (sequential.cc:32-74).
#ifndef PPI_DEVICE_OPS_H
#define PPI_DEVICE_OPS_H

#include <hip/hip_runtime.h>

#include "exact_math.h"
#include "tape.h"

namespace ppi {
namespace dev {

// Float constants of the reference's symbol header / functions.h:19-57.
constexpr float kPI_F = 3.14159274101257f;
constexpr float kPI_2_F = 1.57079637050629f;
constexpr float kE_F = 2.71828174591064f;

__device__ __forceinline__ float b2f(bool b) { return b ? 1.0f : 0.0f; }

template <bool N> __device__ __forceinline__ float f_exp(float x) {
    if constexpr (N) return __expf(x);
    else return exact_exp(x);
}
// The remaining double routines (exp10, exp2, log10, log2, pow, and GAMMA's sinh and
// pow) are called out of line: inlined into every row of the interpreters' unrolled
// handlers, their register needs set the whole kernel's register count, although
// common grammars rarely or never use them.
__device__ __noinline__ float exp10_of_double(float x) { return (float)exp10((double)x); }
__device__ __noinline__ float exp2_of_double(float x) { return (float)exp2((double)x); }
__device__ __noinline__ float log10_of_double(float x) { return (float)log10((double)x); }
__device__ __noinline__ float log2_of_double(float x) { return (float)log2((double)x); }
__device__ __noinline__ float pow_of_double(float x, float y) { return (float)pow((double)x, (double)y); }
__device__ __noinline__ float fmod_float(float x, float y) { return fmodf(x, y); }

template <bool N> __device__ __forceinline__ float f_exp10(float x) {
    if constexpr (N) return __exp10f(x);
    else return exp10_of_double(x);
}
template <bool N> __device__ __forceinline__ float f_exp2(float x) {
    if constexpr (N) return exp2f(x);
    else return exp2_of_double(x);
}
// The device library's double trigonometric routines carry a large-argument
// reduction; inlined into the interpreters' unrolled per-row loops their
// registers add up across rows, so they are only ever called out of line
// (exact_math.h): as the rare fallback of exact_sin / exact_cos, and for tan.
template <bool N> __device__ __forceinline__ float f_sin(float x) {
    if constexpr (N) return __sinf(x);
    else return exact_sin(x);
}
template <bool N> __device__ __forceinline__ float f_cos(float x) {
    if constexpr (N) return __cosf(x);
    else return exact_cos(x);
}
// functions.h:115-129 (SQRT): float(sqrt(double(x))) == sqrtf(x) exactly.
template <bool P, bool N> __device__ __forceinline__ float f_sqrt(float x) {
    if (P && x < 0.0f) return 1.0f;
    return exact_sqrt(x);
}
// functions.h:131-177 (LOG, LOG10, LOG2).
template <bool P, bool N> __device__ __forceinline__ float f_log(float x) {
    if (P && x < 1.0f) return 1.0f;
    if constexpr (N) return __logf(x);
    else return exact_log(x);
}
template <bool P, bool N> __device__ __forceinline__ float f_log10(float x) {
    if (P && x < 1.0f) return 1.0f;
    if constexpr (N) return __log10f(x);
    else return log10_of_double(x);
}
template <bool P, bool N> __device__ __forceinline__ float f_log2(float x) {
    if (P && x < 1.0f) return 1.0f;
    if constexpr (N) return __log2f(x);
    else return log2_of_double(x);
}
// functions.h:99-113 (DIV).
template <bool P, bool N> __device__ __forceinline__ float f_div(float x, float y) {
    if (P && y == 0.0f) return 1.0f;
    if constexpr (N) return __fdividef(x, y);
    else return x / y;
}
// functions.h:179-193 (TAN).
template <bool P, bool N> __device__ __forceinline__ float f_tan(float x) {
    if (P && (fabsf(x) == kPI_2_F || fabsf(x) == 3.0f * kPI_2_F)) return 1.0f;
    if constexpr (N) return __tanf(x);
    else return tan_of_double(x);
}
// interpreter_core:72-75 (POW): pow is outside functions.h's NATIVE set.
__device__ __forceinline__ float f_pow(float x, float y) { return pow_of_double(x, y); }

// GAMMA (interpreter_core:127-129): sinh and pow are double routines in the reference,
// and the last product is formed in double before it is rounded.  Out of line, like the
// other rare double routines.
template <bool P> __device__ __noinline__ float gamma_of_double(float a) {
    const float s1 = f_sqrt<P, false>((float)((double)a * sinh((double)(1.0f / a))));
    const float base = (a / kE_F) * s1;
    const float s2 = f_sqrt<P, false>(2.0f * kPI_F / a);
    return (float)(pow((double)base, (double)a) * (double)s2);
}

// interpreter_core:1-151, unary operators.
template <bool P, bool N> __device__ __forceinline__ float unary(uint32_t u, float a) {
    switch (u) {
    case U_NOT: return b2f(!(a != 0.0f));
    case U_ABS: return fabsf(a);
    case U_POW2: return a * a;
    case U_POW3: return a * a * a;
    case U_POW4: return a * a * a * a;
    case U_POW5: return a * a * a * a * a;
    case U_NEG: return -a;
    case U_ROUND: return roundf(a);
    case U_CEIL: return ceilf(a);
    case U_FLOOR: return floorf(a);
    case U_EXP: return f_exp<N>(a);
    case U_EXP10: return f_exp10<N>(a);
    case U_EXP2: return f_exp2<N>(a);
    case U_SIN: return f_sin<N>(a);
    case U_COS: return f_cos<N>(a);
    case U_STEP: return b2f(a >= 0.0f);
    case U_SIGN: return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f);
    case U_LOGISTIC: return 1.0f / (1.0f + f_exp<N>(-a));
    case U_GAMMA: return gamma_of_double<P>(a);  // sinh, pow and SQRT: the same routines in both modes
    case U_GAUSS: return f_exp<N>(-a * a);
    case U_SQRT: return f_sqrt<P, N>(a);
    case U_LOG: return f_log<P, N>(a);
    case U_LOG10: return f_log10<P, N>(a);
    case U_LOG2: return f_log2<P, N>(a);
    default: return f_tan<P, N>(a);
    }
}

// A unary operator over the R rows of one lane.  exp, log, sin and cos take the
// row-batched exact routines (their table loads issue together, one fallback
// branch for all rows); the others go row by row.
template <int R, bool P, bool N> __device__ __forceinline__ void unary_rows(uint32_t u, float (&v)[R]) {
    if constexpr (!N) {
        if (u == U_EXP) return exact_rows<kFnExp, R>(v);
        if (u == U_SIN) return exact_rows<kFnSin, R>(v);
        if (u == U_COS) return exact_rows<kFnCos, R>(v);
        if (u == U_SQRT) {
            // functions.h:115-129: x < 0 gives 1 under PROTECTED, which is sqrt(1).
            if constexpr (P) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (v[r] < 0.0f) v[r] = 1.0f;
            }
            return exact_rows<kFnSqrt, R>(v);
        }
        if (u == U_LOG) {
            // functions.h:131-147: x < 1 gives 1 under PROTECTED; such rows evaluate log(1).
            bool prot[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                prot[r] = P && v[r] < 1.0f;
                if (prot[r]) v[r] = 1.0f;
            }
            exact_rows<kFnLog, R>(v);
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (prot[r]) v[r] = 1.0f;
            return;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = unary<P, N>(u, v[r]);
}

// fmin / fmax as the reference's x86-64 C library computes them (minsd / maxsd): the
// first operand when it is strictly smaller (larger), otherwise the second; a NaN
// operand yields the other one.  v_min_f32 / v_max_f32 order -0 below +0, so on equal
// zeros they can return the other operand than the reference, and the sign then
// reaches results such as gamma(+-0) = +-inf.
__device__ __forceinline__ float ref_fmin(float x, float y) { return (x < y || isnan(y)) ? x : y; }
__device__ __forceinline__ float ref_fmax(float x, float y) { return (x > y || isnan(y)) ? x : y; }

// Binary operators; `a` is the accumulator, `b` the folded or parked operand.
// Non-R forms compute f(a, b); R forms compute f(b, a).
template <bool P, bool N> __device__ __forceinline__ float binary(uint32_t op, float a, float b) {
    switch (op) {
    case B_ADD: return a + b;
    case B_SUB: return a - b;
    case B_RSUB: return b - a;
    case B_MULT: return a * b;
    case B_DIV: return f_div<P, N>(a, b);
    case B_RDIV: return f_div<P, N>(b, a);
    case B_MEAN: return (a + b) / 2.0f;
    case B_MAX: return ref_fmax(a, b);
    case B_RMAX: return ref_fmax(b, a);
    case B_MIN: return ref_fmin(a, b);
    case B_RMIN: return ref_fmin(b, a);
    case B_MOD: return fmod_float(a, b);
    case B_RMOD: return fmod_float(b, a);
    case B_POW: return f_pow(a, b);
    case B_RPOW: return f_pow(b, a);
    case B_AND: return b2f((a != 0.0f) && (b != 0.0f));
    case B_OR: return b2f((a != 0.0f) || (b != 0.0f));
    case B_XOR: return b2f((a != 0.0f) != (b != 0.0f));
    case B_GT: return b2f(a > b);
    case B_GE: return b2f(a >= b);
    case B_LT: return b2f(a < b);
    case B_LE: return b2f(a <= b);
    case B_EQ: return b2f(a == b);
    default: return b2f(a != b);
    }
}

// ERROR(X,Y) (definitions.h:18) for the built-in error functions.
__device__ __forceinline__ float error_fn(int ef, float X, float Y, const float* cost, int n) {
    switch (ef) {
    case PPI_EF_ABS: return fabsf(X - Y);
    case PPI_EF_SQUARED: return (X - Y) * (X - Y);
    case PPI_EF_NEQ: return b2f(X != Y);
    default: {
        float xx = (ef == PPI_EF_COST_CLAMP) ? fminf(fmaxf(X, 0.0f), (float)(n - 1)) : X;
        bool ok = (xx > -1.0f) && (xx < (float)n) && (Y > -1.0f) && (Y < (float)n);
        return ok ? cost[(int)xx * n + (int)Y] : __builtin_nanf("");
    }
    }
}

}  // namespace dev
}  // namespace ppi

#endif
