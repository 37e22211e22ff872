// kernels_impl.h — HIP kernels of the PPI evaluation engine for gfx950.
//
// Forms (reference: accelerator.cl):
//   FP  (fp_tile_kernel)   row tiles x program chunks.  A wave owns 64*R rows whose
//                          features sit in VGPRs; it walks a chunk of programs whose
//                          instructions are wave-uniform (scalar loads, scalar
//                          branches).  Per-(program, row-block) partials are reduced
//                          by finalize_kernel in a fixed order.  (evaluate_dp, :88-171)
//   PDP (pdp_kernel)       one workgroup per program, waves stride over row tiles,
//                          result written directly.  (evaluate_pdp, :173-268)
//   PP  (pp_kernel)        one lane per program, rows in order; the only form whose
//                          sums follow the reference's row order exactly.  (evaluate_pp, :7-86)
#ifndef PPI_KERNELS_IMPL_H
#define PPI_KERNELS_IMPL_H

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <type_traits>

#include "device_ops.h"
#include "kernels.h"

namespace ppi {
namespace dev {

constexpr int kWaves = 4;              // waves per workgroup in FP / PDP forms
constexpr uint32_t kXcds = 8;          // MI355X: 8 accelerator complex dies, workgroups dealt round robin
constexpr int kBlock = 64 * kWaves;

// Waves per SIMD the register allocator must leave room for (the second
// __launch_bounds__ argument caps VGPRs at 512 / waves) in the FP and batched
// forms.  Measured on gfx950 with the regression workload (ms per launch):
// 4 rows per lane with the 4-slot stack 7 waves (72 VGPRs; the spills left
// sit after exact-math calls and in the non-finite epilogue, and the launch
// is 1.5% faster than at 6 waves, which was 6.8% faster than 5), but 6 waves
// in the persistent batched kernel (its queue state spills on the hot path at
// 7), with the 9-slot stack 5 (96 VGPRs;
// 10% faster than 4 waves without the bound); 2 rows per lane 5 (76 VGPRs: 6
// waves run); 1 row per lane (the 16-slot fallback) takes 4; at 8 rows any
// bound spills, so none is set.  More waves hide more of the scalar dispatch.
// A workgroup is one wave per SIMD, so the waves per SIMD are also the
// workgroups per CU that fit its 160 KB of LDS (feature tile + exact-math
// tables): the bound never asks for more, since spilling for waves that LDS
// cannot hold only costs (32 feature slots at 4 rows: 4 waves).  The PDP form
// keeps its rows' features in VGPRs and is left unbounded.
template <int R, int F, int S, bool Queue = false>
struct MinWaves {
    static constexpr int pref = R == 4 ? (S <= 4 ? (Queue ? 6 : 7) : 5) : (R == 2 ? 5 : (R == 1 ? 4 : 1));
    static constexpr int lds = (F > 0 ? F : 1) * 64 * R * 4 + (int)sizeof(ExactTables) + 64;
    static constexpr int fit = (160 * 1024) / lds;
    static constexpr int value = pref < fit ? pref : (fit > 0 ? fit : 1);
};

template <int K> struct VecT { typedef float type __attribute__((ext_vector_type(K))); };

// The tape and program headers are read-only for the whole launch and every
// read is wave-uniform: reading them through the constant address space lets
// the compiler issue scalar loads (s_load through the scalar cache) instead of
// a vector load + wait per instruction.
template <typename T>
using ConstPtr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ ConstPtr<T> as_const(const T* p) {
    return (ConstPtr<T>)(p);
}

__device__ __forceinline__ ProgMeta load_meta(const ProgMeta* meta, int i) {
    ConstPtr<ProgMeta> q = as_const(meta) + i;
    ProgMeta m;
    m.code_off = q->code_off;
    m.ncode = q->ncode;
    m.size = q->size;
    m.flags = q->flags;
    return m;
}

// ---------------------------------------------------------------------------
// interpreter core: one tape instruction over R rows per lane, wave-uniform.
// ---------------------------------------------------------------------------

// Feature sources: where the attribute operands of a row tile come from.
template <int R, int F>
struct VgprFeat {  // tile held in VGPRs, one F-vector per row (PDP, prediction)
    typedef typename VecT<(F > 0 ? F : 1)>::type VF;
    const VF (&xf)[R];
    __device__ __forceinline__ void get(uint32_t col, float (&b)[R]) const {
#pragma unroll
        for (int r = 0; r < R; ++r) b[r] = xf[r][col];
    }
};
template <int R>
struct GlobalFeat {  // straight from the column-major dataset (more than 32 columns)
    const float* __restrict__ X;
    int64_t ldx;
    const int (&row)[R];
    __device__ __forceinline__ void get(uint32_t col, float (&b)[R]) const {
#pragma unroll
        for (int r = 0; r < R; ++r) b[r] = X[(int64_t)col * ldx + row[r]];
    }
};
template <int R, int F>
struct LdsFeat {  // tile shared by the workgroup in LDS, [col][lane][r] (FP form).  The tile holds the
                  // first ncol = min(input columns, F) columns; the others (input columns beyond F and
                  // the derived columns) come from the dataset through the caches.
    typedef typename VecT<R>::type VR;
    typedef __attribute__((address_space(3))) const VR LdsVR;
    LdsVR* base;  // &tile[0][lane][0]: a lane's R rows of column c are one vector at base[c * 64]
    const float* __restrict__ X0;  // the tile's first row of column 0 (wave-uniform)
    int64_t ldx;
    uint32_t ncol;
    int lane;
    __device__ __forceinline__ void get(uint32_t col, float (&b)[R]) const {
        if (col < ncol) {
            // One typed LDS vector load (ds_read_b128 at R = 4); kept apart from the global
            // branch so the two cannot be merged into a flat load.
            const VR v = base[col * 64];
#pragma unroll
            for (int r = 0; r < R; ++r) b[r] = v[r];
        } else if constexpr (R <= 4) {
            // A buffer load: the column's first row of the tile as a wave-uniform descriptor
            // (third word 0x00020000, the gfx9 raw-buffer format) and a 32-bit lane offset, so no
            // 64-bit address per lane is held in registers (at 4 rows per lane this is what lets
            // 7 waves per SIMD fit without a spill on the hot path).  The descriptor's range is
            // the tile's 64 R rows, all inside the dataset, whose columns are padded to a
            // multiple of 2,048 rows.
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(X0 + (int64_t)col * ldx), (short)0, 64 * R * 4, 0x00020000);
#pragma unroll
            for (int r = 0; r < R; ++r)
                b[r] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, lane * 4, 256 * r, 0));
        } else {
            const float* __restrict__ c = X0 + (int64_t)col * ldx + lane;
#pragma unroll
            for (int r = 0; r < R; ++r) b[r] = c[64 * r];
        }
    }
};

template <int R, int F>
__device__ __forceinline__ auto make_feat(const EvalArgs& A, const float* lds_lane, const int (&row)[R]) {
    if constexpr (F > 0) {
        const int lane = (int)(threadIdx.x & 63);
        const int64_t row0 = __builtin_amdgcn_readfirstlane(row[0] - lane);  // row[r] = row0 + 64 r + lane
        return LdsFeat<R, F>{(typename LdsFeat<R, F>::LdsVR*)lds_lane, A.X + row0, A.ldx, (uint32_t)min(A.ncol, F),
                             lane};
    } else {
        return GlobalFeat<R>{A.X, A.ldx, row};
    }
}

// Register stack of S slots for each of a lane's R rows.  The slot index is
// wave-uniform.  With at least 9 elements per register tuple the backend
// indexes it directly (s_set_gpr_idx_on + one v_mov per row) instead of
// expanding into compare/select chains, so stacks of 9 or more slots keep one
// tuple per row, and smaller stacks keep all rows in one tuple of S*R elements
// (row r of slot k at k*R + r: one index register per access, the row as a
// constant offset).  Slots cost S*R VGPRs, so the smallest stack that holds
// the population's programs frees registers for occupancy.
template <int R, int S, bool Joint = (S < 9)>
struct RegStack {
    static_assert(S >= 9, "per-row register stacks need >= 9 slots for indexed moves");
    typename VecT<S>::type v[R];
    __device__ __forceinline__ float get(int r, uint32_t slot) const { return v[r][slot]; }
    __device__ __forceinline__ void set(int r, uint32_t slot, float x) { v[r][slot] = x; }
};

template <int R, int S>
struct RegStack<R, S, true> {
    static_assert(S * R >= 9, "joint register stacks need >= 9 elements for indexed moves");
    typename VecT<S * R>::type v;
    __device__ __forceinline__ float get(int r, uint32_t slot) const { return v[slot * R + r]; }
    __device__ __forceinline__ void set(int r, uint32_t slot, float x) { v[slot * R + r] = x; }
};

template <int R, int S, bool P, bool N>
struct Core {
    typedef RegStack<R, S> Stk;

    static __device__ __forceinline__ void push(Stk& stk, const float (&acc)[R], uint32_t slot) {
#pragma unroll
        for (int r = 0; r < R; ++r) stk.set(r, slot, acc[r]);
    }

    // Executes instruction (c, a); H_END calls on_end().  The handler's category
    // (bits 5-7) and operator (bits 0-4) select the code through a short
    // wave-uniform branch tree: binary operators first, the six arithmetic
    // operators (+, -, reversed -, *, /, reversed /) ahead of the rest, then the
    // unary operators, then loads, pushes, IF_THEN_ELSE and the program end.
    // The operator is taken as h - (category base) inside each branch: decoding
    // it once up front as h & 31 made the adjoint kernel demote its feature
    // vector to scratch memory (17% slower there).
    template <class FeatT, class EndF>
    static __device__ __forceinline__ void step(uint32_t c, uint32_t a, const FeatT& ft, float (&acc)[R],
                                                Stk& stk, EndF&& on_end) {
        const uint32_t h = c & 0xffu;
        const float cv = __uint_as_float(a);
        if (h >= H_BIN_C) {
            float b[R];
            uint32_t op;
            if (h >= H_BIN_S) {
                op = h - H_BIN_S;
                const uint32_t slot = (c >> 8) & 31u;
#pragma unroll
                for (int r = 0; r < R; ++r) b[r] = stk.get(r, slot);
            } else if (h >= H_BIN_A) {
                op = h - H_BIN_A;
                ft.get(a, b);
            } else {
                op = h - H_BIN_C;
#pragma unroll
                for (int r = 0; r < R; ++r) b[r] = cv;
            }
#define PPI_BCASE(K)                                                           \
    case K:                                                                    \
        _Pragma("unroll") for (int r = 0; r < R; ++r) acc[r] = binary<P, N>(K, acc[r], b[r]); \
        break;
            if (__builtin_expect(op <= B_RDIV, 1)) {
                switch (op) {
                    PPI_BCASE(B_ADD) PPI_BCASE(B_SUB) PPI_BCASE(B_RSUB) PPI_BCASE(B_MULT)
                    PPI_BCASE(B_DIV) PPI_BCASE(B_RDIV)
                default: break;
                }
            } else {
                switch (op) {
                    PPI_BCASE(B_MEAN) PPI_BCASE(B_MAX)
                    PPI_BCASE(B_MIN) PPI_BCASE(B_MOD) PPI_BCASE(B_RMOD) PPI_BCASE(B_POW)
                    PPI_BCASE(B_RPOW) PPI_BCASE(B_AND) PPI_BCASE(B_OR) PPI_BCASE(B_XOR)
                    PPI_BCASE(B_GT) PPI_BCASE(B_GE) PPI_BCASE(B_LT) PPI_BCASE(B_LE)
                    PPI_BCASE(B_EQ) PPI_BCASE(B_NE) PPI_BCASE(B_RMAX) PPI_BCASE(B_RMIN)
                default: break;
                }
            }
#undef PPI_BCASE
        } else if (h >= H_UN) {
            switch (h - H_UN) {
#define PPI_UCASE(K)                                                           \
    case K:                                                                    \
        unary_rows<R, P, N>(K, acc);                                           \
        break;
                PPI_UCASE(U_NOT) PPI_UCASE(U_ABS) PPI_UCASE(U_POW2) PPI_UCASE(U_POW3)
                PPI_UCASE(U_POW4) PPI_UCASE(U_POW5) PPI_UCASE(U_NEG) PPI_UCASE(U_ROUND)
                PPI_UCASE(U_CEIL) PPI_UCASE(U_FLOOR) PPI_UCASE(U_EXP) PPI_UCASE(U_EXP10)
                PPI_UCASE(U_EXP2) PPI_UCASE(U_SIN) PPI_UCASE(U_COS) PPI_UCASE(U_STEP)
                PPI_UCASE(U_SIGN) PPI_UCASE(U_LOGISTIC) PPI_UCASE(U_GAMMA)
                PPI_UCASE(U_GAUSS) PPI_UCASE(U_SQRT) PPI_UCASE(U_LOG)
                PPI_UCASE(U_LOG10) PPI_UCASE(U_LOG2) PPI_UCASE(U_TAN)
#undef PPI_UCASE
            default: break;
            }
        } else {
            switch (h) {
            case H_LOAD_C:
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = cv;
                break;
            case H_LOAD_A:
                ft.get(a, acc);
                break;
            case H_PUSH_C:
                push(stk, acc, (c >> 8) & 31u);
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = cv;
                break;
            case H_PUSH_A:
                push(stk, acc, (c >> 8) & 31u);
                ft.get(a, acc);
                break;
            case H_END:
                on_end(a);  // H_END carries the program's flags in its argument
                break;
            default: {
                // H_TERN: IF_THEN_ELSE (interpreter_core:1-4): (bool)cond ? then : else.
                const uint32_t s1 = (c >> 8) & 31u, s2 = (c >> 13) & 31u;
                const uint32_t wc = (c >> 18) & 3u, wt = (c >> 20) & 3u, we = (c >> 22) & 3u;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float v0 = acc[r], v1 = stk.get(r, s1), v2 = stk.get(r, s2);
                    const float cd = wc == 0 ? v0 : (wc == 1 ? v1 : v2);
                    const float th = wt == 0 ? v0 : (wt == 1 ? v1 : v2);
                    const float el = we == 0 ? v0 : (we == 1 ? v1 : v2);
                    acc[r] = (cd != 0.0f) ? th : el;
                }
                break;
            }
            }
        }
    }

    struct NoEnd {
        __device__ __forceinline__ void operator()(uint32_t) const {}
    };

    // One program from scalar-loaded instructions (one instruction of look-ahead).
    template <class FeatT>
    static __device__ __forceinline__ void run(ConstPtr<Ins> code, uint32_t n, const FeatT& ft, float (&acc)[R],
                                               Stk& stk) {
        run_range(code, 0, n, ft, acc, stk);
    }

    // Instructions [pc0, pc1) of a program; code[pc1] must be readable (look-ahead).
    template <class FeatT>
    static __device__ __forceinline__ void run_range(ConstPtr<Ins> code, uint32_t pc0, uint32_t pc1, const FeatT& ft,
                                                     float (&acc)[R], Stk& stk) {
        uint32_t nc = code[pc0].code, na = code[pc0].arg;
        for (uint32_t pc = pc0; pc < pc1; ++pc) {
            const uint32_t c = nc, a = na;
            nc = code[pc + 1].code;  // code[code_off + ncode] is the program's H_END
            na = code[pc + 1].arg;
            step(c, a, ft, acc, stk, NoEnd{});  // the loop stops before the program's H_END
        }
    }
};

// Index of the calling wave in its workgroup.  threadIdx.x >> 6 is the same in
// every lane of a wave, but the compiler cannot prove it; without the
// readfirstlane every program range, tape offset and loop bound derived from it
// would be treated as divergent (vector loads of the scalar tape, exec-mask
// loops).  Measured: +7.7% on the FP form.
__device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// Copies the exact routines' tables into the workgroup's LDS (exact_math.h).
// Called by every thread at kernel entry, before any thread leaves; a barrier
// must follow before the first operator is evaluated.  NATIVE kernels use the
// hardware approximations and need no tables.
template <bool N>
__device__ __forceinline__ void stage_math() {
    if constexpr (!N) stage_exact_tables();
}

// Cross-lane reductions with DPP (data-parallel primitives: a VALU operand read
// from another lane of the same row, or broadcast from the last lane of a row):
// within each row of 16 lanes, swaps of adjacent lanes, adjacent pairs and
// mirrored halves and rows leave the row's total in every lane; two row
// broadcasts then add row 0 into row 1, row 2 into row 3 and rows 0-1 into rows
// 2-3, so lane 63 holds the wave's total.  Six VALU instructions and no LDS round
// trip; the summation order is fixed, so the result is deterministic.  Lanes a
// broadcast does not write keep `neutral`.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float v, float neutral) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(neutral), __float_as_int(v), CTRL, ROW_MASK,
                                                      0xf, false));
}
enum : int {
    kDppSwap1 = 0xB1,      // quad_perm [1,0,3,2]
    kDppSwap2 = 0x4E,      // quad_perm [2,3,0,1]
    kDppHalfMirror = 0x141,
    kDppMirror = 0x140,
    kDppBcast15 = 0x142,   // lane 15 of each row to the next row
    kDppBcast31 = 0x143,   // lane 31 to rows 2 and 3
};
// The wave's sum, in every lane.
__device__ __forceinline__ float wave_sum(float v) {
    v = v + dpp_f<kDppSwap1, 0xf>(v, 0.0f);
    v = v + dpp_f<kDppSwap2, 0xf>(v, 0.0f);
    v = v + dpp_f<kDppHalfMirror, 0xf>(v, 0.0f);
    v = v + dpp_f<kDppMirror, 0xf>(v, 0.0f);
    v = v + dpp_f<kDppBcast15, 0xa>(v, 0.0f);
    v = v + dpp_f<kDppBcast31, 0xc>(v, 0.0f);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// The wave's maximum (fmaxf), in every lane.
__device__ __forceinline__ float wave_max(float v) {
    const float ninf = -__builtin_inff();
    v = fmaxf(v, dpp_f<kDppSwap1, 0xf>(v, ninf));
    v = fmaxf(v, dpp_f<kDppSwap2, 0xf>(v, ninf));
    v = fmaxf(v, dpp_f<kDppHalfMirror, 0xf>(v, ninf));
    v = fmaxf(v, dpp_f<kDppMirror, 0xf>(v, ninf));
    v = fmaxf(v, dpp_f<kDppBcast15, 0xa>(v, ninf));
    v = fmaxf(v, dpp_f<kDppBcast31, 0xc>(v, ninf));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ uint32_t wave_umin(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_umax(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}

__device__ __forceinline__ bool finite_f(float x) { return !(isinf(x) || isnan(x)); }

// Results that depend on the reference's partition or order of the rows.  The
// engine reduces rows in its own order; the reference does it
//  * PDP (accelerator.cl:173-268): lane L of a pdp_lanes-wide work-group takes
//    rows L, L + pdp_lanes, ...; a non-finite error sets the lane's partial to
//    MAXFLOAT and ends the lane; the partials meet in a tree; inf -> MAXFLOAT.
//    MAXFLOAT or MAXFLOAT/nlin then depends on the lane of an overflowing row
//    (REDUCEMAX: an infinite error*nlin survives only in a lane with no bad row)
//    and on the tree's partial sums next to FLT_MAX;
//  * DP (accelerator.cl:88-171, accelerator.cc:773-805): work-groups of dp_local
//    rows reduce by a tree; the host stops at the first non-finite group result
//    with FLT_MAX (-> FLT_MAX/nlin).  A group sum that overflows thus gives
//    FLT_MAX/nlin where an overflow of the host's sum gives FLT_MAX, and the
//    REDUCEMAX tree, (PE[l+s] > PE[l]) ? PE[l+s] : PE[l], drops a NaN met as the
//    upper operand and keeps one met as the lower, which then also hides the
//    upper operand (an inf row included);
//  * SEQ and PP (sequential.cc:169-179, accelerator.cl:7-86): rows in order, so
//    only a sum near overflow can differ from a reordered one.
// The epilogue settles every case its (total, bad-row record) decides and marks
// the rest by writing kRecheckBits, a NaN no epilogue produces, into raw[p];
// recheck_kernel recomputes those programs in the reference's partition.  The
// bad-row record holds, per program, the minimum and maximum of bad_key over
// the rows with a non-finite error.
constexpr uint32_t kRecheckBits = 0x7fc0e3c1u;

// PDP: the row's reference lane; other strategies: 1 for a NaN error, 0 for an infinite one.
__device__ __forceinline__ uint32_t bad_key(const EvalArgs& A, int row, float e) {
    return A.strategy == PPI_STRATEGY_PDP ? (uint32_t)row % (uint32_t)A.pdp_lanes : (isnan(e) ? 1u : 0u);
}

// A sum of row errors computed in another order than the reference's: the
// epilogue's float total and whether the reference's own rounding decides if its
// sum overflows.  The engine's sum arrives as D, the sum in double of its finite
// float partials, plus ninf partials that overflowed; gb bounds a partial's
// relative rounding error.  For non-negative terms any order's float sum lies
// within a relative (n-1)u of the exact sum S (u = 2^-24; recursive summation is
// the worst case, the reference's lane and group trees are tighter) and
// overflows exactly when it reaches thr = 2^128 - 2^103, so S(1 + nu) < thr is
// finite in every order and S(1 - nu) >= thr infinite in every order; in between
// the program is rechecked.  An overflowed partial's exact sum is at least
// thr(1 - gb).  Cost-matrix errors may be negative, so their sums are trusted
// only far from overflow.
struct SumTotal {
    float total;
    bool undecided;
};
__device__ __forceinline__ SumTotal decide_sum(const EvalArgs& A, double D, double gb, int ninf = 0) {
    const double thr = 0x1.0p128 - 0x1.0p103;
    const double gr = 1.01 * (double)A.nlin * 0x1.0p-24 + 1e-12;
    if (!(D == D)) return {(float)D, true};
    if (A.error_fn >= PPI_EF_COST) return {(float)D, ninf > 0 || !(fabs(D) < 0x1.0p120)};
    if (ninf > 0 || isinf(D)) {
        const double lower = ((double)ninf * thr + (isinf(D) ? 0.0 : D)) * (1.0 - gb);
        return {__builtin_inff(), !(lower * (1.0 - gr) >= thr)};
    }
    if (D * (1.0 + gr) < thr * (1.0 - gb)) return {(float)D, false};
    if (D * (1.0 - gb) * (1.0 - gr) >= thr) return {__builtin_inff(), false};
    return {(float)D, true};
}
constexpr double kUlp = 0x1.0p-24 * 1.01;

// Whether a result needs the recomputation.  undecided: from decide_sum (false for
// maxima and for the row-order sums of the PP and ordered forms under SEQ / PP).
// The error functions other than the cost matrix give non-negative errors, so a
// PDP partial holding MAXFLOAT stays at or above it and two of them sum to inf;
// with a cost matrix (entries of either sign) every PDP program with a bad row is
// recomputed.
__device__ __forceinline__ bool needs_recheck(const EvalArgs& A, float total, bool bad, uint32_t kmin,
                                              uint32_t kmax, bool undecided) {
    if (A.strategy == PPI_STRATEGY_PDP) {
        if (A.reducemax) return bad && !finite_f(total);  // no bad row: a maximum is order-free
        if (!bad) return undecided;
        if (A.error_fn >= PPI_EF_COST) return true;
        // One bad lane: MAXFLOAT + x rounds to inf only for x >= 2^103 (half an ulp
        // of FLT_MAX); every tree partial is at most the total of the other rows.
        return kmin == kmax && !(total < 0x1.0p102f);
    }
    if (A.strategy == PPI_STRATEGY_FP) {
        if (A.reducemax) return bad && kmax == 1u;  // a NaN row
        // An overflow gives FLT_MAX/nlin in a work-group's sum and FLT_MAX in the host's.
        return !bad && (undecided || !finite_f(total));
    }
    return !A.reducemax && !bad && undecided;
}

// Per-program result from (reduced value, bad-row record).  Restates the
// reference's epilogues: sequential.cc:135-189, accelerator.cl:64-84 +
// accelerator.cc:849 (PP), accelerator.cl:157-171 + accelerator.cc:773-805 (DP),
// accelerator.cl:236-267 + accelerator.cc:849 (PDP).
__device__ __forceinline__ void semantics(const EvalArgs& A, float total, uint32_t flags, uint32_t kmin,
                                          uint32_t kmax, int32_t size, float* raw_out, float* vec_out,
                                          bool undecided = false) {
    const float fn = (float)A.nlin;
    const float as = A.alpha * (float)size;
    float raw, vec;
    const bool bad = kmin != 0xffffffffu;
    if (flags & kEmpty) {
        if (A.strategy == PPI_STRATEGY_FP) {
            vec = (A.dp_groups == 1 && !A.reducemax) ? FLT_MAX / fn : FLT_MAX;
            raw = vec;
        } else {
            raw = FLT_MAX;
            vec = (A.strategy == PPI_STRATEGY_SEQ) ? FLT_MAX : raw + as;
        }
    } else if (needs_recheck(A, total, bad, kmin, kmax, undecided)) {
        raw = __uint_as_float(kRecheckBits);
        vec = raw;
    } else if (bad) {
        if (A.strategy == PPI_STRATEGY_PDP) {
            // Settled cases: REDUCEMAX with no infinite row (a lane's MAXFLOAT is the
            // maximum), two or more bad lanes (MAXFLOAT + MAXFLOAT = inf), or one bad
            // lane whose MAXFLOAT absorbs the rest.
            if (A.reducemax) raw = FLT_MAX / fn;
            else raw = (kmin != kmax) ? FLT_MAX : FLT_MAX / fn;
            vec = raw + as;
        } else if (A.strategy == PPI_STRATEGY_PP) {
            raw = FLT_MAX / fn;
            vec = raw + as;
        } else {
            vec = FLT_MAX / fn + as;
            raw = vec;
        }
    } else {
        if (A.strategy == PPI_STRATEGY_PP || A.strategy == PPI_STRATEGY_PDP) {
            raw = finite_f(total) ? total / fn : FLT_MAX;
            vec = raw + as;
        } else {
            vec = finite_f(total) ? total / fn + as : FLT_MAX;
            raw = vec;
        }
    }
    *raw_out = raw;
    *vec_out = vec;
}

// Rows, validity and target of a lane's R rows of a tile starting at row0.
template <int R>
__device__ __forceinline__ void tile_rows(const EvalArgs& A, int64_t row0, int lane, int (&row)[R], float (&y)[R],
                                          bool (&valid)[R]) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
        row[r] = (int)(row0 + r * 64 + lane);
        valid[r] = row[r] < A.nlin;
        y[r] = A.X[(int64_t)A.ycol * A.ldx + row[r]];
    }
}

// Features of a lane's R rows into VGPRs (columns >= ncol read as 0).
template <int R, int F>
__device__ __forceinline__ void tile_features(const EvalArgs& A, const int (&row)[R],
                                              typename VgprFeat<R, F>::VF (&xf)[R]) {
    if constexpr (F > 0) {
#pragma unroll
        for (int f = 0; f < F; ++f) {
            const bool have = f < A.ncol;
#pragma unroll
            for (int r = 0; r < R; ++r) xf[r][f] = have ? A.X[(int64_t)f * A.ldx + row[r]] : 0.0f;
        }
    }
}

// Error of the R rows of one lane: accumulate finite errors, record bad rows.  `full`: every
// row of the tile is a dataset row (wave-uniform).
template <int R>
__device__ __forceinline__ void accumulate(const EvalArgs& A, const float (&acc)[R], const float (&y)[R],
                                           const bool (&valid)[R], const int (&row)[R], bool full, float& part,
                                           uint32_t& kmin, uint32_t& kmax) {
    const float fn = (float)A.nlin;
    float e[R];
#pragma unroll
    for (int r = 0; r < R; ++r) e[r] = error_fn(A.error_fn, acc[r], y[r], A.cost, A.cost_n);
    if (!A.reducemax) {
        // Sums: the rows' errors are added in row order with no per-row test.  A finite result
        // proves every error finite (an inf or NaN error leaves the sum non-finite), and is then
        // the exact rules' result bit for bit; otherwise (a non-finite error, or a sum that
        // overflowed) the rules below recompute it.  Rows past the end, only in the dataset's
        // last tile, add 0.
        float q = part;
        if (full) {
#pragma unroll
            for (int r = 0; r < R; ++r) q = q + e[r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) q = q + (valid[r] ? e[r] : 0.0f);
        }
        if (__builtin_expect(finite_f(q), 1)) {
            part = q;
            return;
        }
    }
    // Branch-free over the rows (a row that is past the end or not finite adds 0, which
    // leaves the partial's bits unchanged: it is never -0 and a maximum of non-negative
    // values is at least 0); the bad-row record is one rarely taken branch.
    bool any_bad = false;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool fin = finite_f(e[r]);
        any_bad = any_bad || (valid[r] && !fin);
        const bool take = valid[r] && fin;
        if (A.reducemax) part = fmaxf(part, take ? e[r] * fn : 0.0f);
        else part += take ? e[r] : 0.0f;
    }
    if (any_bad) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (valid[r] && !finite_f(e[r])) {
                const uint32_t key = bad_key(A, row[r], e[r]);
                kmin = min(kmin, key);
                kmax = max(kmax, key);
            }
        }
    }
}

// Wave-reduce a lane partial and bad-row record; lane 0 stores them for program p.
// (Buffering a wave's partials one program per lane and storing 64 at a time cuts the
// kernel's write traffic by a third, but the lane select costs two more VALU instructions
// per program end than this store and the kernel measured 1.3% slower.)
__device__ __forceinline__ void publish(const EvalArgs& A, int64_t slot, int p, int lane, float part, uint32_t kmin,
                                        uint32_t kmax) {
    part = A.reducemax ? wave_max(part) : wave_sum(part);
    if (__any(kmin != 0xffffffffu)) {
        kmin = wave_umin(kmin);
        kmax = wave_umax(kmax);
        if (lane == 0) {
            atomicMin(&A.bad_lo[p], kmin);
            atomicMax(&A.bad_hi[p], kmax);
        }
    }
    if (lane == 0) A.partial[slot] = part;
}

// ---------------------------------------------------------------------------
// FP form
//
// Workgroup = one tile of 64*R rows x one chunk of programs.  The tile's
// features are staged once in LDS and shared by the 4 waves; each wave takes a
// contiguous quarter of the chunk and streams through its tape.  Instructions
// are fetched from a 64-entry window held across the wave's lanes (lane k holds
// tape entry base+k, read back with v_readlane), refilled by one coalesced load
// per 64 instructions and double-buffered, so no memory latency sits on the
// dispatch path.  Programs end with H_END, where the wave reduces the program's
// error over the tile and stores one partial.
// ---------------------------------------------------------------------------
// Ordered form: the R row errors of one lane for program p, in row order within the
// pass buffer (row block rb - rb0, program p, row r*64 + lane of the tile).
template <int R>
__device__ __forceinline__ void store_row_errors(const EvalArgs& A, int rb, int p, int lane, const float (&acc)[R],
                                                 const float (&y)[R], const bool (&valid)[R]) {
    float* dst = A.rowerr + ((int64_t)(rb - A.rb0) * A.nprog + p) * (64 * R) + lane;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[r * 64] = valid[r] ? error_fn(A.error_fn, acc[r], y[r], A.cost, A.cost_n) : 0.0f;
}

// One FP work item: rows [rb*64R, rb*64R + 64R) of A's dataset x programs [c0, c1)
// of A's population.  Called by every thread of the workgroup; `tile` is the
// workgroup's LDS feature buffer.  The caller separates consecutive items with a
// barrier (the tile is rewritten at the start).
template <int R, int F, int S, bool P, bool N, bool ROWS = false>
__device__ __forceinline__ void fp_tile_body(const EvalArgs& A, int rb, int c0, int c1,
                                             float (*tile)[64][R]) {
    typedef Core<R, S, P, N> C;
    const int lane = threadIdx.x & 63;
    const int wave = wave_index();
    const int64_t row0 = (int64_t)rb * (64 * R);

    if constexpr (F > 0) {
        for (int i = threadIdx.x; i < F * R * 64; i += kBlock) {
            const int f = i / (R * 64), j = i - f * (R * 64);
            tile[f][j & 63][j >> 6] = f < A.ncol ? A.X[(int64_t)f * A.ldx + row0 + j] : 0.0f;
        }
    }
    int row[R];
    float y[R];
    bool valid[R];
    tile_rows<R>(A, row0, lane, row, y, valid);
    const bool full = row0 + 64 * R <= (int64_t)A.nlin;
    __syncthreads();

    // The waves split the programs by tape length, not by count: a wave that finishes early
    // idles its slot until the workgroup ends.  Wave w takes the programs whose first
    // instruction lies in [t_w, t_{w+1}), t_w = g0 + w (g1 - g0) / kWaves (binary searches on
    // the programs' tape offsets, wave-uniform).
    auto first_at = [&](int64_t t) {  // first program in [c0, c1) starting at or after t (c1 if none)
        int lo = c0, hi = c1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int64_t)as_const(A.meta)[mid].code_off < t) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    const ProgMeta mlast = load_meta(A.meta, c1 - 1);
    const int64_t g0 = as_const(A.meta)[c0].code_off;
    const int64_t g1 = (int64_t)mlast.code_off + mlast.ncode + 1;
    const int pa = wave == 0 ? c0 : first_at(g0 + (g1 - g0) * wave / kWaves);
    const int pb = wave == kWaves - 1 ? c1 : first_at(g0 + (g1 - g0) * (wave + 1) / kWaves);
    if (pa >= pb) return;

    const ProgMeta first = load_meta(A.meta, pa);
    const ProgMeta last = load_meta(A.meta, pb - 1);
    const int64_t gbeg = first.code_off;
    const int64_t gend = (int64_t)last.code_off + last.ncode + 1;  // through the last H_END

    typedef typename std::conditional<(F > 0), LdsFeat<R, F>, GlobalFeat<R>>::type FeatT;
    FeatT ft = make_feat<R, F>(A, &tile[0][lane][0], row);

    // Instruction window: lane L holds tape entry blk + L of the current 64-entry
    // block; the next block is loaded when the current one starts, so its latency
    // is hidden behind 64 instructions (the tape is padded so blk + 127 stays in
    // bounds).  blk and k are wave-uniform (scalar registers, scalar loop).
    const Ins* code = A.code;
    int64_t blk = gbeg & ~(int64_t)63;
    uint32_t wc = code[blk + lane].code, wa = code[blk + lane].arg;
    int k = (int)(gbeg - blk);
    int p = pa;
    float acc[R];
    typename C::Stk stk;
    for (; blk < gend; blk += 64) {
        const uint32_t nc = code[blk + 64 + lane].code, na = code[blk + 64 + lane].arg;
        const int kend = (int)min<int64_t>(64, gend - blk);
        for (; k < kend; ++k) {
            const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)wc, k);
            const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)wa, k);
            if ((c & 0xffu) == H_END) {
                // H_END carries the program's flags in its argument.
                if constexpr (ROWS) {
                    if (!(a & kEmpty)) store_row_errors<R>(A, rb, p, lane, acc, y, valid);
                } else {
                    float part = 0.0f;
                    uint32_t kmin = 0xffffffffu, kmax = 0u;
                    if (!(a & kEmpty)) accumulate<R>(A, acc, y, valid, row, full, part, kmin, kmax);
                    publish(A, (int64_t)rb * A.nprog + p, p, lane, part, kmin, kmax);
                }
                ++p;
                continue;
            }
            C::step(c, a, ft, acc, stk, typename C::NoEnd{});
        }
        k = 0;
        wc = nc;
        wa = na;
    }
}

template <int R, int F, int S, bool P, bool N, bool ROWS>
__global__ void __launch_bounds__(kBlock, (MinWaves<R, F, S>::value)) fp_tile_kernel(EvalArgs A) {
    constexpr int FL = F > 0 ? F : 1;
    __shared__ __attribute__((aligned(16))) float tile[FL][64][R];
    stage_math<N>();  // fp_tile_body's barrier follows
    // XCD-aware order.  The hardware hands workgroup i to XCD i mod kXcds; this
    // maps XCD x's workgroups onto a contiguous run of (row tile, chunk) items,
    // chunk fastest, so the chunks of one row tile run at about the same time on
    // one XCD and share its L2 (the tile's derived columns are read by every
    // chunk).  A bijection on [0, tiles * chunks) for any count.
    const uint32_t nch = (uint32_t)((A.nprog + A.chunk - 1) / A.chunk);
    const uint32_t total = gridDim.x, i = blockIdx.x;
    const uint32_t x = i % kXcds, k = i / kXcds, q = total / kXcds, rem = total % kXcds;
    const uint32_t item = x * q + min(x, rem) + k;
    const int rbl = (int)(item / nch);
    const int c0 = (int)(item - (uint32_t)rbl * nch) * A.chunk;
    fp_tile_body<R, F, S, P, N, ROWS>(A, A.rb0 + rbl, c0, min(c0 + A.chunk, A.nprog), tile);
}

// ---------------------------------------------------------------------------
// Batched persistent form: one launch over many (population, dataset) pairs.
// A fixed grid (sized to the device's occupancy) pulls work items — (pair, row
// tile, program range), cut so that every item costs about the same and ordered
// longest first — from an atomic counter until the list is exhausted; every
// workgroup leaves the loop when the counter passes the end.
// ---------------------------------------------------------------------------
template <int R, int F, int S, bool P, bool N>
__global__ void __launch_bounds__(kBlock, (MinWaves<R, F, S, true>::value)) batch_kernel(BatchArgs B) {
    constexpr int FL = F > 0 ? F : 1;
    __shared__ __attribute__((aligned(16))) float tile[FL][64][R];
    __shared__ int item_s;
    stage_math<N>();  // the loop's first barrier follows
    for (;;) {
        __syncthreads();  // previous item's readers are done with the tile and item_s
        if (threadIdx.x == 0) item_s = (int)atomicAdd(B.counter, 1u);
        __syncthreads();
        const int it = item_s;
        if (it >= B.n_items) break;
        const WorkItem w = B.items[it];
        const PairDev pd = B.pairs[w.pair];
        EvalArgs A = B.base;
        A.meta = B.base.meta + pd.prog_base;
        A.X = pd.X;
        A.ldx = pd.ldx;
        A.nlin = pd.nlin;
        A.ncol = pd.ncol;
        A.ycol = pd.ncol - 1;
        A.nprog = pd.nprog;
        A.pdp_lanes = pd.pdp_lanes;
        A.dp_groups = pd.dp_groups;
        A.partial = B.base.partial + pd.partial_base;
        A.bad_lo = B.base.bad_lo + pd.prog_base;
        A.bad_hi = B.base.bad_hi + pd.prog_base;
        fp_tile_body<R, F, S, P, N>(A, w.row_tile, w.p0, w.p1, tile);
    }
}

#ifdef PPI_KERNELS_MISC
// A program's row-tile partials (stride apart) reduced in tile order: the maximum,
// or the sum in double.  A partial is the wave's DPP tree (6 levels) over lane
// sums of at most 8 rows, so every row error passes through at most 14 float
// roundings before the double sum.
__device__ __forceinline__ SumTotal reduce_partials(const EvalArgs& A, const float* partial, int nrb,
                                                    int64_t stride) {
    // The partials are loaded 16 at a time (independent loads in flight together) and taken
    // in tile order, one after another: the order of the additions, and so the bits, do not
    // depend on the batching.
    constexpr int kB = 16;
    if (A.reducemax) {
        float mx = 0.0f;
        int rb = 0;
        for (; rb + kB <= nrb; rb += kB) {
            float v[kB];
#pragma unroll
            for (int k = 0; k < kB; ++k) v[k] = partial[(int64_t)(rb + k) * stride];
#pragma unroll
            for (int k = 0; k < kB; ++k) mx = fmaxf(mx, v[k]);
        }
        for (; rb < nrb; ++rb) mx = fmaxf(mx, partial[(int64_t)rb * stride]);
        return {mx, false};
    }
    double sum = 0.0;
    int ninf = 0;
    auto take = [&](float v) {
        if (isinf(v) && v > 0.0f) ++ninf;
        else sum += (double)v;
    };
    int rb = 0;
    for (; rb + kB <= nrb; rb += kB) {
        float v[kB];
#pragma unroll
        for (int k = 0; k < kB; ++k) v[k] = partial[(int64_t)(rb + k) * stride];
#pragma unroll
        for (int k = 0; k < kB; ++k) take(v[k]);
    }
    for (; rb < nrb; ++rb) take(partial[(int64_t)rb * stride]);
    return decide_sum(A, sum, 14.0 * kUlp + (double)nrb * 0x1.0p-53, ninf);
}

// Per-program epilogue of the batched form: program g belongs to pair pair_of[g].
__global__ void batch_finalize_kernel(BatchArgs B, const int32_t* pair_of, int32_t n_total) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    const PairDev pd = B.pairs[pair_of[g]];
    const int p = g - pd.prog_base;
    EvalArgs A = B.base;
    A.nlin = pd.nlin;
    A.dp_groups = pd.dp_groups;
    A.pdp_lanes = pd.pdp_lanes;
    const SumTotal t = reduce_partials(A, B.base.partial + pd.partial_base + p, pd.nrb, pd.nprog);
    const ProgMeta m = B.base.meta[g];
    semantics(A, t.total, m.flags, B.base.bad_lo[g], B.base.bad_hi[g], m.size, &B.base.raw[g], &B.base.vec[g],
              t.undecided);
}

__global__ void finalize_kernel(EvalArgs A, int nrb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.nprog) return;
    const SumTotal t = reduce_partials(A, A.partial + p, nrb, A.nprog);
    const ProgMeta m = A.meta[p];
    semantics(A, t.total, m.flags, A.bad_lo[p], A.bad_hi[p], m.size, &A.raw[p], &A.vec[p], t.undecided);
}

// Ordered form: one thread per program continues its sum (REDUCEMAX: maximum) over the
// pass's rows in row order, exactly as the sequential loop does (sequential.cc:169-179:
// float accumulation row by row, stop at the first non-finite error), and the last pass
// applies the strategy's epilogue.  A program's row errors are contiguous within each
// row tile; they are read 64 at a time as float4, the next batch in flight while the
// current one is summed, so the serial sum does not wait on memory latency per load.
__device__ __forceinline__ void ordered_take(float e, float fn, bool rmax, float& part, uint32_t& stop) {
    const bool fin = finite_f(e);
    const bool take = !stop && fin;
    const float cand = rmax ? ((e * fn > part) ? e * fn : part) : part + e;
    part = take ? cand : part;
    stop = (stop || !fin) ? 1u : 0u;
}

// Ordered sums: one thread per program, one wave per workgroup (30,000 programs are 469 waves,
// spread over every CU).  Each lane streams its program's row errors of the pass in row order
// (rowerr[tile][program][row of the tile]) in batches of 64 (16 float4), the next batch loading
// while the current one is added.
//
// Full tiles take a fast path: the 64 errors of a batch are added with one instruction each on
// the dependency chain (REDUCEMAX: a compare and a select), with the sequential loop's
// non-finite test moved off the chain: for sums, a sum that is still finite at the end of the
// batch proves that every error in it was finite (an inf or NaN error leaves the sum
// non-finite); REDUCEMAX tests every error's class.  A batch that fails the test (a non-finite
// error, or a sum that overflowed) is added again from its starting value with the exact
// per-row rules (ordered_take), which also sets `stop`.  The last tile of the dataset, when
// partial, is added row by row with the same rules.
constexpr int kOrdV = 16;  // float4 per batch

template <bool RMAX>
__device__ __forceinline__ void ordered_batch(const float4 (&c)[kOrdV], float fn, float& part, uint32_t& stop) {
    float q = part;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < kOrdV; ++k) {
        const float e[4] = {c[k].x, c[k].y, c[k].z, c[k].w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if constexpr (RMAX) {
                bad = bad || !finite_f(e[i]);
                const float x = e[i] * fn;
                q = (x > q) ? x : q;
            } else {
                q = q + e[i];
            }
        }
    }
    if constexpr (!RMAX) bad = !finite_f(q);
    if (__builtin_expect(bad && !stop, 0)) {
        q = part;
        uint32_t st = 0;
#pragma unroll
        for (int k = 0; k < kOrdV; ++k) {
            ordered_take(c[k].x, fn, RMAX, q, st);
            ordered_take(c[k].y, fn, RMAX, q, st);
            ordered_take(c[k].z, fn, RMAX, q, st);
            ordered_take(c[k].w, fn, RMAX, q, st);
        }
        stop = st;
    }
    part = stop ? part : q;
}

template <bool RMAX>
__global__ void __launch_bounds__(64) ordered_sum_kernel(EvalArgs A, int rows_per_block, int last) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = p < A.nprog;
    const int pp = live ? p : A.nprog - 1;  // the tail wave's spare lanes shadow the last program
    const ProgMeta m = A.meta[pp];
    float part = A.osum[pp];
    uint32_t stop = A.ostop[pp] | ((m.flags & kEmpty) ? 1u : 0u);
    const float fn = (float)A.nlin;
    // Full tiles of the pass, then at most one partial tile (the dataset's last).
    const int64_t row_end = (int64_t)(A.rb0 + A.nrb_pass) * rows_per_block;
    const bool partial_last = row_end > (int64_t)A.nlin;
    const int full = A.nrb_pass - (partial_last ? 1 : 0);
    const int bpt = rows_per_block / (4 * kOrdV);  // batches per tile (rows_per_block = 64 R)
    const int nb = full * bpt;
    const int64_t tstride = (int64_t)A.nprog * rows_per_block;  // floats between a program's tiles
    const float* base = A.rowerr + (int64_t)pp * rows_per_block;
    auto batch_ptr = [&](int b) {
        const int t = b / bpt;
        return reinterpret_cast<const float4*>(base + t * tstride + (b - t * bpt) * (4 * kOrdV));
    };
    float4 a[kOrdV], c[kOrdV];
    if (nb > 0) {
        const float4* src = batch_ptr(0);
#pragma unroll
        for (int k = 0; k < kOrdV; ++k) a[k] = src[k];
    }
    for (int b = 0; b < nb; b += 2) {
        // The loads are unconditional (a batch past the end reloads the last one) so that the
        // compiler waits only for the batch about to be added.
        const float4* sc = batch_ptr(min(b + 1, nb - 1));
#pragma unroll
        for (int k = 0; k < kOrdV; ++k) c[k] = sc[k];
        ordered_batch<RMAX>(a, fn, part, stop);
        const float4* sa = batch_ptr(min(b + 2, nb - 1));
#pragma unroll
        for (int k = 0; k < kOrdV; ++k) a[k] = sa[k];
        if (b + 1 < nb) ordered_batch<RMAX>(c, fn, part, stop);
        if (__all(stop)) break;
    }
    if (partial_last && !stop) {
        const int64_t row0 = row_end - rows_per_block;
        const int nrows = (int)((int64_t)A.nlin - row0);
        const float* src = base + (int64_t)full * tstride;
        for (int j = 0; j < nrows && !stop; ++j) ordered_take(src[j], fn, RMAX, part, stop);
    }
    if (!live) return;
    const uint32_t stopped = stop && !(m.flags & kEmpty) ? 1u : A.ostop[p];
    A.osum[p] = part;
    A.ostop[p] = stopped;
    if (last) semantics(A, part, m.flags, stopped ? 0u : 0xffffffffu, 0u, m.size, &A.raw[p], &A.vec[p]);
}

// ---------------------------------------------------------------------------
// Recheck: programs whose epilogue marked them (kRecheckBits in raw[p]) are
// recomputed in the reference's own partition and order of the rows.  Every
// thread of the workgroup runs the program on every pass (rows past the end or
// after a stop are evaluated on row 0 and dropped), so the tape loop stays
// wave-uniform.  Each body ends with thread 0 writing the result; the caller's
// barrier follows.
// ---------------------------------------------------------------------------
constexpr int kRecheckSpan = 16;  // programs per recheck workgroup

// PPI_EF_EXPR: the compiled ERROR(X,Y) (ef.h) on one row, with the host interpreter's rules
// (ef.cc, ef_eval_host): values held in double, a float-typed operation rounded to float after
// it (exact for + - * / and sqrt: double has more than twice float's precision), int
// operations wrapping in 32 bits, conversions to int truncating (INT_MIN out of range), the
// double routines from the device library.  The stack lives in private memory: this runs in
// the recomputation kernels only.
static __device__ __noinline__ double ef_round_to(double v, uint32_t type) {
    if (type == EF_FLOAT) return (double)(float)v;
    if (type == EF_INT) return (v > -2147483649.0 && v < 2147483648.0) ? (double)(int32_t)v : -2147483648.0;
    return v;
}
__device__ __forceinline__ double ef_wrap(int64_t v) { return (double)(int32_t)(uint32_t)(uint64_t)v; }

static __device__ __noinline__ float ef_expr_eval(const EfIns* code, int n, float X, float Y) {
    double st[kEfMaxDepth];
    int sp = 0;
    for (int k = 0; k < n; ++k) {
        const EfIns e = code[k];
        const uint32_t t = e.type;
        switch (e.op) {
        case EF_X: st[sp++] = X; break;
        case EF_Y: st[sp++] = Y; break;
        case EF_IMM: st[sp++] = e.imm; break;
        case EF_CVT: st[sp - 1] = ef_round_to(st[sp - 1], t); break;
        case EF_NEG: st[sp - 1] = t == EF_INT ? ef_wrap(-(int64_t)st[sp - 1]) : -st[sp - 1]; break;
        case EF_NOT: st[sp - 1] = st[sp - 1] == 0.0 ? 1.0 : 0.0; break;
        case EF_FABS: {
            const double v = st[sp - 1];
            st[sp - 1] = t == EF_INT ? ef_wrap(v < 0.0 ? -(int64_t)v : (int64_t)v) : fabs(v);
            break;
        }
        case EF_SQRT: st[sp - 1] = ef_round_to(sqrt(st[sp - 1]), t); break;
        case EF_EXP: st[sp - 1] = ef_round_to(exp(st[sp - 1]), t); break;
        case EF_LOG: st[sp - 1] = ef_round_to(log(st[sp - 1]), t); break;
        case EF_SIN: st[sp - 1] = ef_round_to(sin(st[sp - 1]), t); break;
        case EF_COS: st[sp - 1] = ef_round_to(cos(st[sp - 1]), t); break;
        case EF_TAN: st[sp - 1] = ef_round_to(tan(st[sp - 1]), t); break;
        case EF_FLOOR: st[sp - 1] = floor(st[sp - 1]); break;
        case EF_CEIL: st[sp - 1] = ceil(st[sp - 1]); break;
        case EF_ROUND: st[sp - 1] = round(st[sp - 1]); break;
        case EF_TRUNC: st[sp - 1] = trunc(st[sp - 1]); break;
        case EF_SEL: {
            const double b = st[--sp], a = st[--sp], c = st[--sp];
            st[sp++] = c != 0.0 ? a : b;
            break;
        }
        default: {
            const double b = st[--sp], a = st[--sp];
            double r = 0.0;
            switch (e.op) {
            case EF_ADD: r = t == EF_INT ? ef_wrap((int64_t)a + (int64_t)b) : ef_round_to(a + b, t); break;
            case EF_SUB: r = t == EF_INT ? ef_wrap((int64_t)a - (int64_t)b) : ef_round_to(a - b, t); break;
            case EF_MUL: r = t == EF_INT ? ef_wrap((int64_t)a * (int64_t)b) : ef_round_to(a * b, t); break;
            case EF_DIV:
                if (t == EF_INT) r = b == 0.0 ? 0.0 : ef_wrap((int64_t)a / (int64_t)b);
                else r = ef_round_to(a / b, t);
                break;
            case EF_MOD: r = b == 0.0 ? 0.0 : ef_wrap((int64_t)a % (int64_t)b); break;
            case EF_LT: r = a < b; break;
            case EF_GT: r = a > b; break;
            case EF_LE: r = a <= b; break;
            case EF_GE: r = a >= b; break;
            case EF_EQ: r = a == b; break;
            case EF_NE: r = a != b; break;
            case EF_AND: r = (a != 0.0) && (b != 0.0); break;
            case EF_OR: r = (a != 0.0) || (b != 0.0); break;
            case EF_POW: r = ef_round_to(pow(a, b), t); break;
            case EF_FMIN: r = fmin(a, b); break;
            case EF_FMAX: r = fmax(a, b); break;
            default: break;
            }
            st[sp++] = r;
            break;
        }
        }
    }
    return (float)st[sp - 1];
}

// Error of program m at row n.
template <bool P, bool N>
__device__ __forceinline__ float row_error(const EvalArgs& A, const ProgMeta& m, int n) {
    typedef Core<1, kMaxSlots, P, N> C;
    int row[1] = {n};
    float acc[1];
    typename C::Stk stk;
    C::run(as_const(A.code) + m.code_off, m.ncode, GlobalFeat<1>{A.X, A.ldx, row}, acc, stk);
    const float y = A.X[(int64_t)A.ycol * A.ldx + n];
    if (A.error_fn == PPI_EF_EXPR) return ef_expr_eval(A.ef, A.ef_n, acc[0], y);
    return error_fn(A.error_fn, acc[0], y, A.cost, A.cost_n);
}

// Sources of a program's row errors for the partition-exact paths below: computed on the spot
// (the recheck kernels), or read from a buffer that deep_rows_kernel filled (deep_partition_kernel).
template <bool P, bool N>
struct RowErrCompute {
    const EvalArgs& A;
    const ProgMeta& m;
    __device__ __forceinline__ float operator()(int n) const { return row_error<P, N>(A, m, n); }
};
struct RowErrBuffer {
    const float* e;
    __device__ __forceinline__ float operator()(int n) const { return e[n]; }
};

// The reference's work-group reduction over PE[0, width) (accelerator.cl:157-166,
// 251-260): strides from half the next power of two of width down to 1; at stride
// s, lane l < s with l + s < width takes PE[l + s] (sum, or REDUCEMAX's
// (PE[l+s] > PE[l]) ? PE[l+s] : PE[l]).  Lanes below s read PE[l + s] (>= s, not
// written at this stride).  Entered after a barrier; leaves the result in PE[0].
__device__ __forceinline__ void reference_tree(float* PE, int width, bool rmax) {
    int np2 = 1;
    while (np2 < width) np2 <<= 1;
    for (int s = np2 >> 1; s > 0; s >>= 1) {
        for (int l = (int)threadIdx.x; l < s; l += kBlock) {
            if (l + s < width) PE[l] = rmax ? ((PE[l + s] > PE[l]) ? PE[l + s] : PE[l]) : PE[l] + PE[l + s];
        }
        __syncthreads();
    }
}

// PDP (accelerator.cl:196-267): thread t plays lanes t, t + kBlock, ... of the
// pdp_lanes-wide work-group; a lane sums (REDUCEMAX: maximum of error*nlin) its
// rows in order and stops at its first non-finite error with MAXFLOAT; an
// infinite or NaN tree result becomes MAXFLOAT.  The host adds alpha*size
// (accelerator.cc:849).
template <class ErrF>
__device__ __forceinline__ void pdp_recheck(const EvalArgs& A, const ProgMeta& m, float* raw_out, float* vec_out,
                                            float* PE, const ErrF& err) {
    const int lanes = A.pdp_lanes;
    const int passes = (A.nlin + lanes - 1) / lanes;
    const float fn = (float)A.nlin;
    for (int l0 = 0; l0 < lanes; l0 += kBlock) {
        const int l = l0 + (int)threadIdx.x;
        float pe = 0.0f;
        bool stop = l >= lanes;
        for (int j = 0; j < passes; ++j) {
            const int n = j * lanes + l;
            const bool live = !stop && n < A.nlin;
            const float e = err(live ? n : 0);
            if (live) {
                if (!finite_f(e)) {
                    pe = FLT_MAX;
                    stop = true;
                } else if (A.reducemax) {
                    pe = (e * fn > pe) ? e * fn : pe;
                } else {
                    pe += e;
                }
            }
        }
        if (l < lanes) PE[l] = pe;
    }
    __syncthreads();
    reference_tree(PE, lanes, A.reducemax != 0);
    if (threadIdx.x == 0) {
        const float raw = finite_f(PE[0]) ? PE[0] / fn : FLT_MAX;
        *raw_out = raw;
        *vec_out = raw + A.alpha * (float)m.size;
    }
}

// DP (accelerator.cl:113-167, accelerator.cc:773-805): work-group g holds the
// error of row g*dp_local + l in lane l (0 past the last row) and reduces by the
// tree; the host walks the groups in order, stops at the first non-finite group
// result with FLT_MAX, and otherwise sums them (REDUCEMAX: maximum of
// result*nlin); a non-finite total becomes FLT_MAX, else total/nlin + alpha*size.
template <class ErrF>
__device__ __forceinline__ void dp_recheck(const EvalArgs& A, const ProgMeta& m, float* raw_out, float* vec_out,
                                           float* PE, int* stop_s, const ErrF& err) {
    const int width = A.dp_local;
    const int groups = (A.nlin + width - 1) / width;
    const float fn = (float)A.nlin;
    float sum = 0.0f;  // thread 0's
    for (int g = 0; g < groups; ++g) {
        for (int l0 = 0; l0 < width; l0 += kBlock) {
            const int l = l0 + (int)threadIdx.x;
            const int n = g * width + l;
            const bool live = l < width && n < A.nlin;
            const float e = err(live ? n : 0);
            if (l < width) PE[l] = live ? e : 0.0f;
        }
        __syncthreads();
        reference_tree(PE, width, A.reducemax != 0);
        if (threadIdx.x == 0) {
            const float v = PE[0];
            if (!finite_f(v)) {
                sum = FLT_MAX;
                *stop_s = 1;
            } else if (A.reducemax) {
                sum = (v * fn > sum) ? v * fn : sum;
            } else {
                sum += v;
            }
        }
        __syncthreads();  // PE[0] read and the stop published before PE is rewritten
        if (*stop_s) break;
    }
    if (threadIdx.x == 0) {
        const float r = finite_f(sum) ? sum / fn + A.alpha * (float)m.size : FLT_MAX;
        *raw_out = r;
        *vec_out = r;
    }
}

// SEQ / PP (sequential.cc:169-179, accelerator.cl:40-84): rows in order, the
// errors of kBlock rows at a time staged in LDS and taken by thread 0.
template <class ErrF>
__device__ __forceinline__ void ordered_recheck(const EvalArgs& A, const ProgMeta& m, float* raw_out,
                                                float* vec_out, float* PE, int* stop_s, const ErrF& err) {
    const float fn = (float)A.nlin;
    const bool rmax = A.reducemax != 0;
    float part = 0.0f;  // thread 0's
    uint32_t stop = 0;
    for (int n0 = 0; n0 < A.nlin; n0 += kBlock) {
        const int n = n0 + (int)threadIdx.x;
        const float e = err(n < A.nlin ? n : 0);
        PE[threadIdx.x] = e;
        __syncthreads();
        if (threadIdx.x == 0) {
            // 64 errors per batch, read as float4 before the serial takes.
            const int cnt = min(kBlock, A.nlin - n0);
            const float4* v4 = reinterpret_cast<const float4*>(PE);
            for (int b = 0; b < cnt && !stop; b += 64) {
                float4 q[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) q[k] = v4[b / 4 + k];
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const int j = b + 4 * k;
                    if (j + 0 < cnt) ordered_take(q[k].x, fn, rmax, part, stop);
                    if (j + 1 < cnt) ordered_take(q[k].y, fn, rmax, part, stop);
                    if (j + 2 < cnt) ordered_take(q[k].z, fn, rmax, part, stop);
                    if (j + 3 < cnt) ordered_take(q[k].w, fn, rmax, part, stop);
                }
            }
            *stop_s = (int)stop;
        }
        __syncthreads();
        if (*stop_s) break;
    }
    if (threadIdx.x == 0) semantics(A, part, m.flags, stop ? 0u : 0xffffffffu, 0u, m.size, raw_out, vec_out);
}

// The sum of every row error in double, in every thread (relative error about
// nlin/256 * 2^-53 against the exact sum).  red: kBlock doubles of LDS.
template <bool P, bool N>
__device__ __forceinline__ double sum_rows_double(const EvalArgs& A, const ProgMeta& m, double* red) {
    double s = 0.0;
    for (int n0 = 0; n0 < A.nlin; n0 += kBlock) {
        const int n = n0 + (int)threadIdx.x;
        const float e = row_error<P, N>(A, m, n < A.nlin ? n : 0);
        if (n < A.nlin) s += (double)e;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();  // every thread has read red[0] before the buffer is reused
    return r;
}

// The same sum from the program's row-tile partials (pp[rb * A.nprog], A.part_nrb tiles of
// A.part_rows rows): finite partials as they are, and the rows of every tile whose float partial
// overflowed recomputed in double, so a program with one huge tile costs that tile's rows, not a
// pass over the dataset.  The tiles are visited in order by every thread (the same address, so
// the branch is uniform) and each thread's share is fixed, so the result does not vary between
// runs.  Relative error against the exact sum: a partial's at most 14 float roundings (as in
// reduce_partials) plus the double additions.
template <bool P, bool N>
__device__ __forceinline__ double sum_tiles_double(const EvalArgs& A, const ProgMeta& m, const float* pp,
                                                   double* red) {
    double s = 0.0;
    for (int rb = (int)threadIdx.x; rb < A.part_nrb; rb += kBlock) {
        const float v = pp[(int64_t)rb * A.nprog];
        if (!isinf(v)) s += (double)v;
    }
    for (int rb = 0; rb < A.part_nrb; ++rb) {
        if (!isinf(pp[(int64_t)rb * A.nprog])) continue;
        const int r1 = min((rb + 1) * A.part_rows, A.nlin);
        for (int n = rb * A.part_rows + (int)threadIdx.x; n < r1; n += kBlock) s += (double)row_error<P, N>(A, m, n);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// A program marked for its sum is first settled from the double sum when that sum
// is clear of float overflow by more than any order's rounding (and, under DP,
// finite: an overflow's sentinel depends on the work-groups); the rest take the
// partition-exact path.  bad: the bad-row record where the form keeps one (FP and
// batched forms); elsewhere a bad row makes the double sum NaN or infinite, which
// also sends the program on.  pp: the program's first row-tile partial where the
// form keeps them (A.part_nrb > 0), else null.  A program still undecided then takes
// the deep list when A has one with room (A.deep_cap > 0: datasets long enough that one
// workgroup per program would dominate), else the partition-exact path right here.
// The partition-exact result of program m, from the row errors err(n), for the strategy.
template <class ErrF>
__device__ __forceinline__ void partition_result(const EvalArgs& A, const ProgMeta& m, float* raw_out,
                                                 float* vec_out, float* PE, int* stop_s, const ErrF& err) {
    if (A.strategy == PPI_STRATEGY_PDP) pdp_recheck(A, m, raw_out, vec_out, PE, err);
    else if (A.strategy == PPI_STRATEGY_FP) dp_recheck(A, m, raw_out, vec_out, PE, stop_s, err);
    else ordered_recheck(A, m, raw_out, vec_out, PE, stop_s, err);
}

template <bool P, bool N>
__device__ __forceinline__ void recheck_program(const EvalArgs& A, const ProgMeta& m, int id, bool bad,
                                                float* raw_out, float* vec_out, float* PE, int* stop_s,
                                                int* deep_s, const float* pp) {
    if (!bad && !A.reducemax) {
        double* red = reinterpret_cast<double*>(PE);
        const SumTotal t =
            pp ? decide_sum(A, sum_tiles_double<P, N>(A, m, pp, red),
                            14.0 * kUlp + (double)(A.part_nrb + A.nlin) * 0x1.0p-53)
               : decide_sum(A, sum_rows_double<P, N>(A, m, red), 1e-9);
        if (!t.undecided && (A.strategy != PPI_STRATEGY_FP || finite_f(t.total))) {
            if (threadIdx.x == 0) semantics(A, t.total, m.flags, 0xffffffffu, 0u, m.size, raw_out, vec_out);
            return;
        }
    }
    if (A.deep_cap > 0) {
        // Queued for the deep kernels when the list has room (id: where raw_out and vec_out
        // point in A.raw and A.vec).
        if (threadIdx.x == 0) {
            const uint32_t k = atomicAdd(A.deep_count, 1u);
            if (k < (uint32_t)A.deep_cap) A.deep_prog[k] = id;
            *deep_s = k < (uint32_t)A.deep_cap ? 1 : 0;
        }
        __syncthreads();
        const bool queued = *deep_s != 0;
        __syncthreads();  // every thread has read the flag before it is written again
        if (queued) return;
    }
    partition_result(A, m, raw_out, vec_out, PE, stop_s, RowErrCompute<P, N>{A, m});
}

__device__ __forceinline__ bool marked(const float* raw) { return __float_as_uint(*raw) == kRecheckBits; }

// Which of programs [p0, p0 + kRecheckSpan) (below n) are marked, as a bit mask: lane q of
// wave 0 reads program p0 + q's mark, so the span costs one load round trip.  Called by every
// thread of the workgroup (a barrier inside); the marks are read before any recheck rewrites
// its program's result.
__device__ __forceinline__ uint32_t span_marks(const float* raw, int p0, int n, uint32_t* marks_s) {
    static_assert(kRecheckSpan <= 32, "the span's marks fit one 32-bit mask");
    if (threadIdx.x < 64) {
        const int p = p0 + (int)threadIdx.x;
        const bool m = threadIdx.x < (unsigned)kRecheckSpan && p < n && marked(&raw[p]);
        const uint64_t bal = __ballot(m);
        if (threadIdx.x == 0) *marks_s = (uint32_t)bal;
    }
    __syncthreads();
    return *marks_s;
}

// Workgroup b looks at programs [b*kRecheckSpan, (b+1)*kRecheckSpan); the mark is
// read by every thread from the same address, so the branches are uniform.
template <bool P, bool N>
__global__ void __launch_bounds__(kBlock) recheck_kernel(EvalArgs A) {
    __shared__ __attribute__((aligned(16))) float PE[kPdpMaxLanes];
    __shared__ int stop_s, deep_s;
    __shared__ uint32_t marks_s;
    bool staged = false;
    uint32_t marks = span_marks(A.raw, (int)blockIdx.x * kRecheckSpan, A.nprog, &marks_s);
    while (marks) {
        const int q = __builtin_ctz(marks);
        marks &= marks - 1u;
        const int p = (int)blockIdx.x * kRecheckSpan + q;
        if (!staged) {
            stage_math<N>();
            staged = true;
        }
        if (threadIdx.x == 0) stop_s = 0;
        __syncthreads();
        recheck_program<P, N>(A, load_meta(A.meta, p), p, A.bad_lo[p] != 0xffffffffu, &A.raw[p], &A.vec[p], PE,
                              &stop_s, &deep_s, A.part_nrb > 0 ? A.partial + p : nullptr);
        __syncthreads();
    }
}

// PPI_EF_EXPR: empty programs get their result here (as the main kernels' epilogue gives it);
// every other program is marked for the recomputation with a bad-row record, so that
// recheck_program takes the reference's own partition and row order for the strategy.
__global__ void ef_prepare_kernel(EvalArgs A) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.nprog) return;
    const ProgMeta m = A.meta[p];
    if (m.flags & kEmpty) {
        semantics(A, 0.0f, m.flags, 0xffffffffu, 0u, m.size, &A.raw[p], &A.vec[p]);
    } else {
        A.raw[p] = __uint_as_float(kRecheckBits);
        A.bad_lo[p] = 0u;
    }
}

// The pair fields of EvalArgs for pair pd (the rest as in `base`).
__device__ __forceinline__ EvalArgs pair_args(const EvalArgs& base, const PairDev& pd) {
    EvalArgs A = base;
    A.X = pd.X;
    A.ldx = pd.ldx;
    A.nlin = pd.nlin;
    A.ncol = pd.ncol;
    A.ycol = pd.ncol - 1;
    A.nprog = pd.nprog;
    A.pdp_lanes = pd.pdp_lanes;
    A.dp_groups = pd.dp_groups;
    A.dp_local = pd.dp_local;
    A.part_nrb = pd.nrb;
    A.part_rows = pd.rows_per_tile;
    return A;
}

// The recheck for the batched form: program g belongs to pair pair_of[g].
template <bool P, bool N>
__global__ void __launch_bounds__(kBlock) batch_recheck_kernel(BatchArgs B, const int32_t* pair_of,
                                                              int32_t n_total) {
    __shared__ __attribute__((aligned(16))) float PE[kPdpMaxLanes];
    __shared__ int stop_s, deep_s;
    __shared__ uint32_t marks_s;
    bool staged = false;
    uint32_t marks = span_marks(B.base.raw, (int)blockIdx.x * kRecheckSpan, n_total, &marks_s);
    while (marks) {
        const int q = __builtin_ctz(marks);
        marks &= marks - 1u;
        const int g = (int)blockIdx.x * kRecheckSpan + q;
        if (!staged) {
            stage_math<N>();
            staged = true;
        }
        if (threadIdx.x == 0) stop_s = 0;
        __syncthreads();
        const PairDev pd = B.pairs[pair_of[g]];
        const EvalArgs A = pair_args(B.base, pd);
        recheck_program<P, N>(A, load_meta(B.base.meta, g), g, B.base.bad_lo[g] != 0xffffffffu, &B.base.raw[g],
                              &B.base.vec[g], PE, &stop_s, &deep_s,
                              B.base.partial + pd.partial_base + (g - pd.prog_base));
        __syncthreads();
    }
}

// Deep list, step 1: the row errors of every queued program, spread over the whole grid.  Item u
// = (entry u / chunks, rows [c * kBlock, (c + 1) * kBlock) with c = u % chunks); a thread per row.
template <bool P, bool N>
__global__ void __launch_bounds__(kBlock) deep_rows_kernel(DeepArgs D) {
    const uint32_t queued = *D.base.deep_count;
    const int cnt = (int)min(queued, (uint32_t)D.base.deep_cap);
    if (cnt == 0) return;
    stage_math<N>();
    __syncthreads();
    const int64_t chunks = ((int64_t)D.max_rows + kBlock - 1) / kBlock;
    for (int64_t u = blockIdx.x; u < cnt * chunks; u += gridDim.x) {
        const int k = (int)(u / chunks);
        const int n = (int)(u - (int64_t)k * chunks) * kBlock + (int)threadIdx.x;
        const int id = D.base.deep_prog[k];
        const EvalArgs A = D.pairs ? pair_args(D.base, D.pairs[D.pair_of[id]]) : D.base;
        if (n < A.nlin) D.err[(int64_t)k * D.stride + n] = row_error<P, N>(A, load_meta(A.meta, id), n);
    }
}

// SEQ / PP from a program's row errors e[0, nlin) (sequential.cc:169-179 and ordered_take's
// rules).  With a bad row the result does not depend on the sum, and with none a REDUCEMAX
// maximum does not depend on the order, so both are found by all threads; only the plain sum is
// taken row by row, by lane 0 of wave 0, from LDS blocks that waves 1-3 stage one block ahead.
constexpr int kDeepStage = 2048;  // floats per staged block
__device__ __forceinline__ void ordered_from_buffer(const EvalArgs& A, const ProgMeta& m, const float* e,
                                                    float* raw_out, float* vec_out,
                                                    float (*stage)[kDeepStage], float* red) {
    const float fn = (float)A.nlin;
    const int tid = (int)threadIdx.x;
    bool bad = false;
    float mx = 0.0f;
    for (int n = tid; n < A.nlin; n += kBlock) {
        const float v = e[n];
        bad = bad || !finite_f(v);
        const float c = v * fn;
        mx = (c > mx) ? c : mx;
    }
    if (__syncthreads_or(bad)) {
        if (tid == 0) semantics(A, 0.0f, m.flags, 0u, 0u, m.size, raw_out, vec_out);
        return;
    }
    if (A.reducemax) {
        red[tid] = mx;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (tid < w) red[tid] = fmaxf(red[tid], red[tid + w]);
            __syncthreads();
        }
        if (tid == 0) semantics(A, red[0], m.flags, 0xffffffffu, 0u, m.size, raw_out, vec_out);
        return;
    }
    const int nblk = (A.nlin + kDeepStage - 1) / kDeepStage;
    for (int i = tid; i < kDeepStage; i += kBlock) stage[0][i] = i < A.nlin ? e[i] : 0.0f;
    __syncthreads();
    float part = 0.0f;  // lane 0 of wave 0's
    for (int b = 0; b < nblk; ++b) {
        if (tid >= 64) {
            const int64_t base = (int64_t)(b + 1) * kDeepStage;
            for (int i = tid - 64; i < kDeepStage; i += kBlock - 64)
                stage[(b + 1) & 1][i] = base + i < A.nlin ? e[base + i] : 0.0f;
        } else if (tid == 0) {
            const float* s = stage[b & 1];
            const int cnt = min(kDeepStage, A.nlin - b * kDeepStage);
            const float4* s4 = reinterpret_cast<const float4*>(s);
            int i = 0;
#pragma unroll 8
            for (; i < cnt / 4; ++i) {
                const float4 q = s4[i];
                part += q.x;
                part += q.y;
                part += q.z;
                part += q.w;
            }
            for (i *= 4; i < cnt; ++i) part += s[i];
        }
        __syncthreads();
    }
    if (tid == 0) semantics(A, part, m.flags, 0xffffffffu, 0u, m.size, raw_out, vec_out);
}

// Deep list, step 2: workgroup k reduces entry k's row errors in the reference's partition and
// order for the strategy and writes the program's result.
__global__ void __launch_bounds__(kBlock) deep_partition_kernel(DeepArgs D) {
    __shared__ __attribute__((aligned(16))) float PE[kPdpMaxLanes];
    __shared__ __attribute__((aligned(16))) float stage[2][kDeepStage];
    __shared__ int stop_s;
    const int k = blockIdx.x;
    const uint32_t queued = *D.base.deep_count;
    if (k >= (int)min(queued, (uint32_t)D.base.deep_cap)) return;
    const int id = D.base.deep_prog[k];
    const EvalArgs A = D.pairs ? pair_args(D.base, D.pairs[D.pair_of[id]]) : D.base;
    const float* err = D.err + (int64_t)k * D.stride;
    const ProgMeta m = load_meta(A.meta, id);
    if (A.strategy != PPI_STRATEGY_PDP && A.strategy != PPI_STRATEGY_FP) {
        ordered_from_buffer(A, m, err, &A.raw[id], &A.vec[id], stage, PE);
        return;
    }
    if (threadIdx.x == 0) stop_s = 0;
    __syncthreads();
    partition_result(A, m, &A.raw[id], &A.vec[id], PE, &stop_s, RowErrBuffer{err});
}
#endif  // PPI_KERNELS_MISC

// ---------------------------------------------------------------------------
// PDP form: one workgroup per program.
// ---------------------------------------------------------------------------
template <int R, int F, int S, bool P, bool N>
__global__ void __launch_bounds__(kBlock) pdp_kernel(EvalArgs A) {
    typedef Core<R, S, P, N> C;
    __shared__ float red[kWaves];
    __shared__ uint32_t rlo[kWaves], rhi[kWaves];
    stage_math<N>();
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = wave_index();
    const int p = blockIdx.x;
    const ProgMeta m = load_meta(A.meta, p);
    float part = 0.0f;
    uint32_t kmin = 0xffffffffu, kmax = 0u;
    if (!(m.flags & kEmpty)) {
        const int64_t step = (int64_t)kWaves * 64 * R;
        for (int64_t row0 = (int64_t)wave * 64 * R; row0 < A.nlin; row0 += step) {
            int row[R];
            float y[R];
            bool valid[R];
            tile_rows<R>(A, row0, lane, row, y, valid);
            float acc[R];
            typename C::Stk stk;
            if constexpr (F > 0) {
                typename VgprFeat<R, F>::VF xf[R];
                tile_features<R, F>(A, row, xf);
                C::run(as_const(A.code) + m.code_off, m.ncode, VgprFeat<R, F>{xf}, acc, stk);
            } else {
                C::run(as_const(A.code) + m.code_off, m.ncode, GlobalFeat<R>{A.X, A.ldx, row}, acc, stk);
            }
            accumulate<R>(A, acc, y, valid, row, row0 + 64 * R <= (int64_t)A.nlin, part, kmin, kmax);
        }
    }
    part = A.reducemax ? wave_max(part) : wave_sum(part);
    kmin = wave_umin(kmin);
    kmax = wave_umax(kmax);
    if (lane == 0) { red[wave] = part; rlo[wave] = kmin; rhi[wave] = kmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float mx = red[0];
        double sum = 0.0;
        int ninf = 0;
        uint32_t lo = rlo[0], hi = rhi[0];
        for (int w = 0; w < kWaves; ++w) {
            mx = fmaxf(mx, red[w]);
            if (isinf(red[w]) && red[w] > 0.0f) ++ninf;
            else sum += (double)red[w];
            lo = min(lo, rlo[w]);
            hi = max(hi, rhi[w]);
        }
        // A lane sums its rows serially (at most nlin / 64 of them), then the wave's DPP tree (6 levels).
        const SumTotal t = A.reducemax ? SumTotal{mx, false}
                                       : decide_sum(A, sum, ((double)A.nlin / 64.0 + 8.0) * kUlp + 1e-15, ninf);
        semantics(A, t.total, m.flags, lo, hi, m.size, &A.raw[p], &A.vec[p], t.undecided);
    }
}

// ---------------------------------------------------------------------------
// PP form: one lane per program, rows in order (divergent dispatch).  A
// workgroup is one wave (64 programs, longest first, so more workgroups than
// CUs at the regression size); its programs' instructions are staged in LDS
// when they fit (kPPTape entries), so each lane's per-row instruction fetch is
// an LDS read rather than a divergent global one.
// ---------------------------------------------------------------------------
constexpr int kPPBlock = 64;
constexpr int kPPTape = 4096;  // instructions staged per workgroup (32 KB)

template <bool P, bool N>
__global__ void __launch_bounds__(kPPBlock) pp_kernel(EvalArgs A) {
    __shared__ float stk[kMaxSlots][kPPBlock];
    __shared__ Ins tape[kPPTape];
    __shared__ uint32_t scan[kPPBlock];
    stage_math<N>();
    const int t = blockIdx.x * kPPBlock + threadIdx.x;
    const bool live = t < A.nprog;
    const int p = live ? (A.order ? A.order[t] : t) : 0;
    const ProgMeta m = live ? A.meta[p] : ProgMeta{0u, 0u, 0, kEmpty};
    // Exclusive prefix sum of the instruction counts (Hillis-Steele over the workgroup).
    const uint32_t nins = (m.flags & kEmpty) ? 0u : m.ncode;
    uint32_t incl = nins;
    scan[threadIdx.x] = incl;
    __syncthreads();
    for (int d = 1; d < kPPBlock; d <<= 1) {
        const uint32_t add = (int)threadIdx.x >= d ? scan[threadIdx.x - d] : 0u;
        __syncthreads();
        incl += add;
        scan[threadIdx.x] = incl;
        __syncthreads();
    }
    const uint32_t total = scan[kPPBlock - 1], off = incl - nins;
    const bool staged = total <= (uint32_t)kPPTape;  // the same in every lane
    if (staged)
        for (uint32_t k = 0; k < nins; ++k) tape[off + k] = A.code[m.code_off + k];
    __syncthreads();  // tape and math tables staged
    if (!live) return;
    float part = 0.0f;
    double dsum = 0.0;  // DP / PDP sums: the reference adds in its own tree order, not row order
    uint32_t kmin = 0xffffffffu, kmax = 0u;
    // SEQ / PP stop at the first bad row (sequential.cc:173, accelerator.cl:64); DP sums
    // only need to know that one exists; the DP maximum keeps looking for a NaN row;
    // PDP needs every bad lane.
    const bool in_order = A.strategy == PPI_STRATEGY_SEQ || A.strategy == PPI_STRATEGY_PP;
    const bool stop_at_bad = in_order || (A.strategy == PPI_STRATEGY_FP && !A.reducemax);
    const float fn = (float)A.nlin;
    if (!(m.flags & kEmpty)) {
        const Ins* code = staged ? &tape[off] : A.code + m.code_off;
        for (int row = 0; row < A.nlin; ++row) {
            float acc = 0.0f;
            for (uint32_t pc = 0; pc < m.ncode; ++pc) {
                const uint32_t c = code[pc].code, a = code[pc].arg;
                const uint32_t h = c & 0xffu;
                const float cv = __uint_as_float(a);
                if (h >= H_BIN_C) {
                    float b;
                    uint32_t op;
                    if (h < H_BIN_A) { op = h - H_BIN_C; b = cv; }
                    else if (h < H_BIN_S) { op = h - H_BIN_A; b = A.X[(int64_t)a * A.ldx + row]; }
                    else { op = h - H_BIN_S; b = stk[(c >> 8) & 31u][threadIdx.x]; }
                    acc = binary<P, N>(op, acc, b);
                } else if (h >= H_UN) {
                    acc = unary<P, N>(h - H_UN, acc);
                } else if (h == H_LOAD_C) {
                    acc = cv;
                } else if (h == H_LOAD_A) {
                    acc = A.X[(int64_t)a * A.ldx + row];
                } else if (h == H_PUSH_C || h == H_PUSH_A) {
                    stk[(c >> 8) & 31u][threadIdx.x] = acc;
                    acc = (h == H_PUSH_C) ? cv : A.X[(int64_t)a * A.ldx + row];
                } else {
                    const float v1 = stk[(c >> 8) & 31u][threadIdx.x], v2 = stk[(c >> 13) & 31u][threadIdx.x];
                    const uint32_t wc = (c >> 18) & 3u, wt = (c >> 20) & 3u, we = (c >> 22) & 3u;
                    const float cd = wc == 0 ? acc : (wc == 1 ? v1 : v2);
                    const float th = wt == 0 ? acc : (wt == 1 ? v1 : v2);
                    const float el = we == 0 ? acc : (we == 1 ? v1 : v2);
                    acc = (cd != 0.0f) ? th : el;
                }
            }
            const float e = error_fn(A.error_fn, acc, A.X[(int64_t)A.ycol * A.ldx + row], A.cost, A.cost_n);
            if (!finite_f(e)) {
                const uint32_t key = bad_key(A, row, e);
                kmin = min(kmin, key);
                kmax = max(kmax, key);
                if (stop_at_bad || (A.strategy == PPI_STRATEGY_FP && kmax == 1u)) break;
            } else if (A.reducemax) {
                part = (e * fn > part) ? e * fn : part;
            } else if (in_order) {
                part += e;
            } else {
                dsum += (double)e;
            }
        }
    }
    // SEQ / PP: the row-order float sum is the reference's own.  DP / PDP: a row-order
    // float sum can stray from the reference's tree-ordered one by up to nlin roundings
    // when a few rows dominate, so the rows are summed in double (relative error about
    // nlin * 2^-53) and the float total is that sum rounded.
    const SumTotal tot = (in_order || A.reducemax) ? SumTotal{part, false}
                                                   : decide_sum(A, dsum, (double)A.nlin * 0x1.0p-52);
    semantics(A, tot.total, m.flags, kmin, kmax, m.size, &A.raw[p], &A.vec[p], tot.undecided);
}

// ---------------------------------------------------------------------------
// prediction (ppp -pred): one program, one row per lane.
// ---------------------------------------------------------------------------
template <int F, int S, bool P, bool N>
__global__ void __launch_bounds__(kBlock) predict_kernel(EvalArgs A, float* out) {
    typedef Core<1, S, P, N> C;
    stage_math<N>();
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * kWaves + wave_index()) * 64;
    int row[1];
    float y[1];
    bool valid[1];
    tile_rows<1>(A, row0, lane, row, y, valid);
    const ProgMeta m = load_meta(A.meta, 0);
    float acc[1] = {__builtin_nanf("")};
    typename C::Stk stk;
    if (!(m.flags & kEmpty)) {
        if constexpr (F > 0) {
            typename VgprFeat<1, F>::VF xf[1];
            tile_features<1, F>(A, row, xf);
            C::run(as_const(A.code) + m.code_off, m.ncode, VgprFeat<1, F>{xf}, acc, stk);
        } else {
            C::run(as_const(A.code) + m.code_off, m.ncode, GlobalFeat<1>{A.X, A.ldx, row}, acc, stk);
        }
    }
    if (valid[0]) out[row[0]] = acc[0];
}

#ifdef PPI_KERNELS_MISC
// Derived columns: X[(first + j) * ldx + row] = unary(op_j, X[src_j * ldx + row]) for
// j = 0 .. count-1 and every row of the padded column, with the interpreters' own
// operator code.  A thread owns one row and computes the columns in order, so a
// column whose source is an earlier derived column reads its finished value.
template <bool P, bool N>
__global__ void derive_kernel(float* X, int64_t ldx, const uint32_t* ops, const int32_t* srcs, int32_t first,
                              int32_t count) {
    stage_math<N>();
    __syncthreads();
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= ldx) return;
    for (int32_t j = 0; j < count; ++j)
        X[(int64_t)(first + j) * ldx + row] = unary<P, N>(ops[j], X[(int64_t)srcs[j] * ldx + row]);
}

// Folded constant subtrees (FoldSet): one wave per fold program, run by the
// interpreters' own step code (wave-uniform, like a program of the FP form), so
// a fold's value has the bits the unfolded program would compute on any row.
// Fold programs read no attribute; the accessor below is never called.
struct NoFeat {
    __device__ __forceinline__ void get(uint32_t, float (&b)[1]) const { b[0] = __builtin_nanf(""); }
};

template <bool P, bool N>
__global__ void __launch_bounds__(kBlock) fold_kernel(const ProgMeta* meta, const Ins* code, int32_t n, float* out) {
    typedef Core<1, kMaxSlots, P, N> C;
    stage_math<N>();
    __syncthreads();
    const int64_t f = (int64_t)blockIdx.x * kWaves + wave_index();
    if (f >= n) return;  // whole wave
    const ProgMeta m = load_meta(meta, (int32_t)f);
    float acc[1] = {__builtin_nanf("")};
    typename C::Stk stk;
    C::run(as_const(code) + m.code_off, m.ncode, NoFeat{}, acc, stk);
    if ((threadIdx.x & 63) == 0) out[f] = acc[0];
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
#endif  // PPI_KERNELS_MISC

}  // namespace dev

// ---------------------------------------------------------------------------
// launcher templates
// ---------------------------------------------------------------------------
namespace launch {

template <int R, int F, int S, bool ROWS>
hipError_t launch_fp_rt(const EvalArgs& A, bool P, bool N, hipStream_t st) {
    const int64_t rows_per_block = fp_rows_per_block(R);
    // All row tiles, or the ordered form's pass of nrb_pass tiles from rb0.
    const int64_t tiles = ROWS ? A.nrb_pass : (A.nlin + rows_per_block - 1) / rows_per_block;
    const int64_t chunks = (A.nprog + A.chunk - 1) / A.chunk;
    dim3 grid((unsigned)(tiles * chunks));  // fp_tile_kernel maps the 1-D index onto (tile, chunk)
    if (P && N) hipLaunchKernelGGL((dev::fp_tile_kernel<R, F, S, true, true, ROWS>), grid, dim3(dev::kBlock), 0, st, A);
    else if (P) hipLaunchKernelGGL((dev::fp_tile_kernel<R, F, S, true, false, ROWS>), grid, dim3(dev::kBlock), 0, st, A);
    else if (N) hipLaunchKernelGGL((dev::fp_tile_kernel<R, F, S, false, true, ROWS>), grid, dim3(dev::kBlock), 0, st, A);
    else hipLaunchKernelGGL((dev::fp_tile_kernel<R, F, S, false, false, ROWS>), grid, dim3(dev::kBlock), 0, st, A);
    return hipGetLastError();
}

template <int R, int F, int S>
hipError_t launch_fp_t(const EvalArgs& A, bool P, bool N, bool rows, hipStream_t st) {
    return rows ? launch_fp_rt<R, F, S, true>(A, P, N, st) : launch_fp_rt<R, F, S, false>(A, P, N, st);
}

template <int R, int F, int S>
hipError_t launch_batch_t(const BatchArgs& B, int grid, bool P, bool N, hipStream_t st) {
    if (P && N) hipLaunchKernelGGL((dev::batch_kernel<R, F, S, true, true>), dim3(grid), dim3(dev::kBlock), 0, st, B);
    else if (P) hipLaunchKernelGGL((dev::batch_kernel<R, F, S, true, false>), dim3(grid), dim3(dev::kBlock), 0, st, B);
    else if (N) hipLaunchKernelGGL((dev::batch_kernel<R, F, S, false, true>), dim3(grid), dim3(dev::kBlock), 0, st, B);
    else hipLaunchKernelGGL((dev::batch_kernel<R, F, S, false, false>), dim3(grid), dim3(dev::kBlock), 0, st, B);
    return hipGetLastError();
}

template <int R, int F, int S>
int batch_occupancy_t(bool P, bool N) {
    int n = 0;
    hipError_t e;
    if (P && N) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::batch_kernel<R, F, S, true, true>, dev::kBlock, 0);
    else if (P) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::batch_kernel<R, F, S, true, false>, dev::kBlock, 0);
    else if (N) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::batch_kernel<R, F, S, false, true>, dev::kBlock, 0);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::batch_kernel<R, F, S, false, false>, dev::kBlock, 0);
    return e == hipSuccess ? n : 0;
}

template <int R, int F, int S>
hipError_t launch_pdp_t(const EvalArgs& A, bool P, bool N, hipStream_t st) {
    dim3 grid((unsigned)A.nprog);
    if (P && N) hipLaunchKernelGGL((dev::pdp_kernel<R, F, S, true, true>), grid, dim3(dev::kBlock), 0, st, A);
    else if (P) hipLaunchKernelGGL((dev::pdp_kernel<R, F, S, true, false>), grid, dim3(dev::kBlock), 0, st, A);
    else if (N) hipLaunchKernelGGL((dev::pdp_kernel<R, F, S, false, true>), grid, dim3(dev::kBlock), 0, st, A);
    else hipLaunchKernelGGL((dev::pdp_kernel<R, F, S, false, false>), grid, dim3(dev::kBlock), 0, st, A);
    return hipGetLastError();
}

template <int F, int S>
hipError_t launch_predict_t(const EvalArgs& A, bool P, bool N, float* out, hipStream_t st) {
    dim3 grid((unsigned)((A.nlin + dev::kBlock - 1) / dev::kBlock));
    if (P && N) hipLaunchKernelGGL((dev::predict_kernel<F, S, true, true>), grid, dim3(dev::kBlock), 0, st, A, out);
    else if (P) hipLaunchKernelGGL((dev::predict_kernel<F, S, true, false>), grid, dim3(dev::kBlock), 0, st, A, out);
    else if (N) hipLaunchKernelGGL((dev::predict_kernel<F, S, false, true>), grid, dim3(dev::kBlock), 0, st, A, out);
    else hipLaunchKernelGGL((dev::predict_kernel<F, S, false, false>), grid, dim3(dev::kBlock), 0, st, A, out);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace ppi

#endif
