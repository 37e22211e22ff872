// adjoint_impl.h — reverse-mode adjoint kernels for gfx950: d(fitness)/d(T_CONST).
//
// fitness = sum_rows EF(pred, y) / nlin + alpha*size (sequential.cc:170-189), so
// d fitness / d c = sum_rows EF'(pred, y) * d pred / d c / nlin.  Waves of a
// persistent grid take (job, row range) units from a queue counter; a wave walks
// its unit's range 64*R rows at a time (R = 1 or 2 rows per lane):
//
//   for each segment s of the job:
//       forward pass over the program's accumulator tape (up to the last
//       instruction that stores a partial of segment s), plain steps between
//       the segment's stores, which also write their partials;
//       s == 0: seed the root adjoint with EF'(pred, y) (0 on rows past nlin)
//               and add the row's error to the fitness sum;
//       the segment's reverse instructions: adj[dst] = adj[src] * partial, or
//       g[k] += adj[src] * partial for a constant k (a zero adjoint gives 0
//       whatever the partial).
//
// Tape, store list and reverse tape are fetched through 64-entry windows held
// across the wave's lanes (v_readlane per instruction, no memory latency on the
// dispatch path); a job whose program, store list and reverse tape each fit one
// window keeps them in registers for the whole unit.  The partial store is either
// VGPRs (16 / R per row: long programs re-run the forward pass per segment —
// recomputation) or the wave's own LDS block (32 / R per row: most programs need
// one forward pass — storage).  Adjoint slots and the per-lane gradient
// accumulators live in VGPRs; wave-uniform slot indices use the backend's indexed
// register moves.  Sums over a range are wave-reduced once per unit; ranges are
// summed in a fixed order by adj_grad_kernel.
#ifndef PPI_ADJOINT_IMPL_H
#define PPI_ADJOINT_IMPL_H

#include "kernels_impl.h"

namespace ppi {
namespace dev {

constexpr float kLN10_F = 2.30258509299404568402f;
constexpr float kLN2_F = 0.69314718055994530942f;

// d GAMMA / da (interpreter_core:127-129 differentiated), in double.
template <bool P>
__device__ __noinline__ float dgamma(float a, float z) {
    const double x = a, e = (double)kE_F, pi = (double)kPI_F;
    const double in1 = x * sinh(1.0 / x), in2 = 2.0 * pi / x;
    const bool p1 = P && in1 < 0.0, p2 = P && in2 < 0.0;
    const double s1 = p1 ? 1.0 : sqrt(in1), s2 = p2 ? 1.0 : sqrt(in2);
    const double b = (x / e) * s1;
    const double ds1 = p1 ? 0.0 : 0.5 / s1 * (sinh(1.0 / x) - cosh(1.0 / x) / x);
    const double ds2 = p2 ? 0.0 : 0.5 / s2 * (-2.0 * pi / (x * x));
    const double db = s1 / e + (x / e) * ds1;
    return (float)((double)z * (log(b) + x * db / b + ds2 / s2));
}

// d unary(u, a) / da, given z = unary(u, a).  Protected operators have derivative
// 0 where the protection replaced the value (functions.h:115-193).
template <bool P, bool N>
__device__ __forceinline__ float dunary(uint32_t u, float a, float z) {
    switch (u) {
    case U_ABS: return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f);
    case U_POW2: return 2.0f * a;
    case U_POW3: return 3.0f * a * a;
    case U_POW4: return 4.0f * a * a * a;
    case U_POW5: return 5.0f * a * a * a * a;
    case U_NEG: return -1.0f;
    case U_EXP: return z;
    case U_EXP10: return z * kLN10_F;
    case U_EXP2: return z * kLN2_F;
    case U_SIN:
        if constexpr (N) return __cosf(a);
        else return cosf(a);
    case U_COS:
        if constexpr (N) return -__sinf(a);
        else return -sinf(a);
    case U_LOGISTIC: return z * (1.0f - z);
    case U_GAMMA: return dgamma<P>(a, z);
    case U_GAUSS: return -2.0f * a * z;
    case U_SQRT: return (P && a < 0.0f) ? 0.0f : 0.5f / z;
    case U_LOG: return (P && a < 1.0f) ? 0.0f : 1.0f / a;
    case U_LOG10: return (P && a < 1.0f) ? 0.0f : 1.0f / (a * kLN10_F);
    case U_LOG2: return (P && a < 1.0f) ? 0.0f : 1.0f / (a * kLN2_F);
    case U_TAN: return (P && (fabsf(a) == kPI_2_F || fabsf(a) == 3.0f * kPI_2_F)) ? 0.0f : 1.0f + z * z;
    default: return 0.0f;  // NOT, ROUND, CEIL, FLOOR, STEP, SIGN: cut, never stored
    }
}

// Partials of z = binary(op, a, b) with respect to a (the accumulator) and b.
// MAX/MIN pass the derivative to the operand they return, the left one on ties
// or when the right one is NaN (fmax/fmin semantics).
template <bool P, bool N>
__device__ __forceinline__ void dbinary(uint32_t op, float a, float b, float z, float& da, float& db) {
    switch (op) {
    case B_ADD: da = 1.0f; db = 1.0f; return;
    case B_SUB: da = 1.0f; db = -1.0f; return;
    case B_RSUB: da = -1.0f; db = 1.0f; return;
    case B_MULT: da = b; db = a; return;
    case B_MEAN: da = 0.5f; db = 0.5f; return;
    case B_DIV: {
        const bool zr = P && b == 0.0f;
        da = zr ? 0.0f : 1.0f / b;
        db = zr ? 0.0f : -z / b;
        return;
    }
    case B_RDIV: {
        const bool zr = P && a == 0.0f;
        db = zr ? 0.0f : 1.0f / a;
        da = zr ? 0.0f : -z / a;
        return;
    }
    case B_MAX: da = (isnan(b) || a >= b) ? 1.0f : 0.0f; db = 1.0f - da; return;
    case B_RMAX: db = (isnan(a) || b >= a) ? 1.0f : 0.0f; da = 1.0f - db; return;
    case B_MIN: da = (isnan(b) || a <= b) ? 1.0f : 0.0f; db = 1.0f - da; return;
    case B_RMIN: db = (isnan(a) || b <= a) ? 1.0f : 0.0f; da = 1.0f - db; return;
    case B_MOD: da = 1.0f; db = -truncf(a / b); return;
    case B_RMOD: db = 1.0f; da = -truncf(b / a); return;
    case B_POW: da = b * powf(a, b - 1.0f); db = a > 0.0f ? z * logf(a) : 0.0f; return;
    case B_RPOW: db = a * powf(b, a - 1.0f); da = b > 0.0f ? z * logf(b) : 0.0f; return;
    default: da = 0.0f; db = 0.0f; return;
    }
}

// Partial stores: slot s (< B) of row r.
template <int R, int B>
struct PStoreV {  // VGPRs, one tuple of B*R (>= 16: indexed moves, see RegStack)
    typedef typename VecT<B * R>::type V;
    V v;
    __device__ __forceinline__ void set(uint32_t s, int r, float x) { v[s * R + r] = x; }
    __device__ __forceinline__ float get(uint32_t s, int r) const { return v[s * R + r]; }
};
template <int R>
struct PStoreL {  // this lane's column of the wave's [B][R][64] LDS block
    float* base;
    __device__ __forceinline__ void set(uint32_t s, int r, float x) { base[(s * R + r) * 64] = x; }
    __device__ __forceinline__ float get(uint32_t s, int r) const { return base[(s * R + r) * 64]; }
};

// Partials per row: 16 / R in registers (PS 0), 32 / R in LDS (PS 1, 8 KB per wave).
template <int PS, int R> struct StoreCfg { static constexpr int B = (PS == 1 ? 32 : 16) / R; };

__device__ __forceinline__ AdjJob load_job(const AdjJob* jobs, int i) {
    ConstPtr<AdjJob> q = as_const(jobs) + i;
    AdjJob j;
    j.code_off = q->code_off;
    j.ncode = q->ncode;
    j.seg_off = q->seg_off;
    j.nseg = q->nseg;
    j.prog = q->prog;
    j.const_base = q->const_base;
    j.nconst = q->nconst;
    j.flags = q->flags;
    j.cost = 0;
    j.cols = q->cols;
    j.pad_[0] = j.pad_[1] = 0;
    return j;
}

template <int R, int S, int PS, bool P, bool N>
struct AdjCore {
    typedef Core<R, S, P, N> C;
    typedef typename C::Stk Stk;
    static constexpr int B = StoreCfg<PS, R>::B;
    static constexpr uint32_t kNone = 0xffffffffu;

    // One instruction that stores partials: sa / sb = store slot of the partial
    // with respect to the accumulator / the other operand (IF_THEN_ELSE: the
    // then / else selector), kNone = not stored in this segment.
    template <class FeatT, class PSt>
    static __device__ __forceinline__ void step_rec(uint32_t c, uint32_t a, uint32_t sa, uint32_t sb,
                                                    const FeatT& ft, float (&acc)[R], Stk& stk, PSt& ps) {
        const uint32_t h = c & 0xffu;
        if (h == H_TERN) {
            const uint32_t s1 = (c >> 8) & 31u, s2 = (c >> 13) & 31u;
            const uint32_t wc = (c >> 18) & 3u, wt = (c >> 20) & 3u, we = (c >> 22) & 3u;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float v0 = acc[r], v1 = stk.get(r, s1), v2 = stk.get(r, s2);
                const float cd = wc == 0 ? v0 : (wc == 1 ? v1 : v2);
                const float th = wt == 0 ? v0 : (wt == 1 ? v1 : v2);
                const float el = we == 0 ? v0 : (we == 1 ? v1 : v2);
                const bool t = cd != 0.0f;
                if (sa != kNone) ps.set(sa, r, t ? 1.0f : 0.0f);
                if (sb != kNone) ps.set(sb, r, t ? 0.0f : 1.0f);
                acc[r] = t ? th : el;
            }
        } else if (h >= H_BIN_C) {
            const float cv = __uint_as_float(a);
            float b[R];
            uint32_t op;
            if (h < H_BIN_A) {
                op = h - H_BIN_C;
#pragma unroll
                for (int r = 0; r < R; ++r) b[r] = cv;
            } else if (h < H_BIN_S) {
                op = h - H_BIN_A;
                ft.get(a, b);
            } else {
                op = h - H_BIN_S;
                const uint32_t slot = (c >> 8) & 31u;
#pragma unroll
                for (int r = 0; r < R; ++r) b[r] = stk.get(r, slot);
            }
            switch (op) {
#define PPI_DBCASE(K)                                                                  \
    case K:                                                                            \
        _Pragma("unroll") for (int r = 0; r < R; ++r) {                                \
            const float z = binary<P, N>(K, acc[r], b[r]);                             \
            float da, db;                                                              \
            dbinary<P, N>(K, acc[r], b[r], z, da, db);                                 \
            if (sa != kNone) ps.set(sa, r, da);                                        \
            if (sb != kNone) ps.set(sb, r, db);                                        \
            acc[r] = z;                                                                \
        }                                                                              \
        break;
                PPI_DBCASE(B_MULT) PPI_DBCASE(B_DIV) PPI_DBCASE(B_RDIV) PPI_DBCASE(B_MAX)
                PPI_DBCASE(B_MIN) PPI_DBCASE(B_MOD) PPI_DBCASE(B_RMOD) PPI_DBCASE(B_POW)
                PPI_DBCASE(B_RPOW) PPI_DBCASE(B_RMAX) PPI_DBCASE(B_RMIN)
#undef PPI_DBCASE
            default:  // operators whose stored partials are never requested
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = binary<P, N>(op, acc[r], b[r]);
                break;
            }
        } else {
            // H_UN (leaves never store partials)
            switch (h - H_UN) {
#define PPI_DUCASE(K)                                                                  \
    case K:                                                                            \
        _Pragma("unroll") for (int r = 0; r < R; ++r) {                                \
            const float z = unary<P, N>(K, acc[r]);                                    \
            ps.set(sa, r, dunary<P, N>(K, acc[r], z));                                 \
            acc[r] = z;                                                                \
        }                                                                              \
        break;
                PPI_DUCASE(U_ABS) PPI_DUCASE(U_POW2) PPI_DUCASE(U_POW3) PPI_DUCASE(U_POW4)
                PPI_DUCASE(U_POW5) PPI_DUCASE(U_EXP) PPI_DUCASE(U_EXP10) PPI_DUCASE(U_EXP2)
                PPI_DUCASE(U_SIN) PPI_DUCASE(U_COS) PPI_DUCASE(U_LOGISTIC) PPI_DUCASE(U_GAMMA)
                PPI_DUCASE(U_GAUSS) PPI_DUCASE(U_SQRT) PPI_DUCASE(U_LOG) PPI_DUCASE(U_LOG10)
                PPI_DUCASE(U_LOG2) PPI_DUCASE(U_TAN)
#undef PPI_DUCASE
            default:
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = unary<P, N>(h - H_UN, acc[r]);
                break;
            }
        }
    }

    // Forward pass over the program's instructions [0, n), fetched through a
    // 64-entry window across the wave's lanes as in the FP form (lane L holds
    // tape entry blk + L, read back with v_readlane; the next block is loaded
    // when the current one starts).  A program shorter than 64 instructions is
    // one window loaded once per unit (res: wc0 / wa0 hold code[c0 + lane]).
    // The segment's stores are entries [ks, ks_end) of the store window st_pc /
    // st_sl, in increasing pc: instruction pc = st_pc[ks] also writes its
    // partials.  code is the whole tape, the program starts at c0; the tape is
    // padded so that blk + 127 stays in bounds.
    template <class FeatT, class PSt>
    static __device__ __forceinline__ void forward(const Ins* code, uint32_t c0, uint32_t n, bool res, uint32_t wc0,
                                                   uint32_t wa0, uint32_t st_pc, uint32_t st_sl, uint32_t ks,
                                                   uint32_t ks_end, const FeatT& ft, float (&acc)[R], Stk& stk,
                                                   PSt& ps) {
        const int lane = threadIdx.x & 63;
        const uint32_t gend = c0 + n;
        uint32_t blk, wc, wa;
        if (res) {
            blk = c0;
            wc = wc0;
            wa = wa0;
        } else {
            blk = c0 & ~63u;
            wc = code[blk + lane].code;
            wa = code[blk + lane].arg;
        }
        uint32_t k = c0 - blk;
        uint32_t next = ks < ks_end ? c0 + (uint32_t)__builtin_amdgcn_readlane((int)st_pc, (int)ks) : 0xffffffffu;
        for (; blk < gend; blk += 64) {
            uint32_t nc = 0, na = 0;
            if (!res) {
                nc = code[blk + 64 + lane].code;
                na = code[blk + 64 + lane].arg;
            }
            const uint32_t kend = min(64u, gend - blk);
            for (; k < kend; ++k) {
                const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)wc, (int)k);
                const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)wa, (int)k);
                if (blk + k != next) {
                    C::step(c, a, ft, acc, stk, typename C::NoEnd{});
                    continue;
                }
                // slot + 1 per side, 0 = not stored: the subtraction gives kNone for 0.
                const uint32_t sl = (uint32_t)__builtin_amdgcn_readlane((int)st_sl, (int)ks);
                step_rec(c, a, (sl & 0xffffu) - 1u, (sl >> 16) - 1u, ft, acc, stk, ps);
                ++ks;
                next = ks < ks_end ? c0 + (uint32_t)__builtin_amdgcn_readlane((int)st_pc, (int)ks) : 0xffffffffu;
            }
            k = 0;
            wc = nc;
            wa = na;
        }
    }
};

template <int R, int F>
struct AdjFeat {
    typedef typename std::conditional<(F > 0), VgprFeat<R, F>, GlobalFeat<R>>::type type;
};

// R rows per lane, F features in VGPRs (0: read from the dataset), S stack
// slots, AS adjoint slots (both register stacks: RegStack's layout), PS
// partial store (0: VGPRs, 1: LDS).
// One (job, row range) unit: unit u is range u / njobs of job order[u % njobs].
// Called by a whole wave; pl is the lane's column of the wave's LDS partials.
template <int R, int F, int S, int AS, int PS, bool P, bool N>
__device__ __forceinline__ void adjoint_unit(const AdjArgs& A, int64_t u, float* pl, int lane) {
    typedef AdjCore<R, S, PS, P, N> AC;
    typedef typename AC::Stk Stk;
    constexpr int B = AC::B;
    static_assert(B < 64, "a segment's stores must fit one window");
    static_assert(AS <= kAdjSlots, "adjoint slots");
    typedef RegStack<R, AS> VA;
    typedef typename VecT<kAdjMaxConsts>::type VG;
    const int rng = (int)(u / A.njobs);
    const int ji = as_const(A.order)[u - (int64_t)rng * A.njobs];
    const AdjJob J = load_job(A.jobs, ji);
    const int64_t rbeg = (int64_t)rng * A.range_rows;
    const int64_t rend = min((int64_t)A.nlin, rbeg + A.range_rows);

    typename std::conditional<PS == 1, PStoreL<R>, PStoreV<R, B>>::type ps;
    if constexpr (PS == 1) ps.base = pl;
    else (void)pl;

    VG g = 0.0f;
    float part = 0.0f;
    bool bad = false;
    if (!(J.flags & kEmpty)) {
        const RevIns* rv = A.rev;
        const Ins* code = A.code;
        const AdjStore* stores = A.stores;
        // Windows that serve every row block of the unit: the program when it is
        // shorter than 64 instructions, the job's reverse tape and store list
        // (all segments, consecutive) when each holds at most 64 entries.  The
        // reverse tape and the store list are padded by 128 and 64 entries.
        const uint32_t c0 = J.code_off;
        const bool fres = J.ncode < 64u;
        const uint32_t fwc = fres ? code[c0 + lane].code : 0u, fwa = fres ? code[c0 + lane].arg : 0u;
        ConstPtr<AdjSeg> sg0 = as_const(A.segs) + J.seg_off;
        ConstPtr<AdjSeg> sgl = sg0 + (J.nseg - 1);
        const uint32_t rb0 = sg0->rev_begin, st0 = sg0->st_begin;
        const bool rres = sgl->rev_end - rb0 <= 64u, sres = sgl->st_end - st0 <= 64u;
        const uint32_t rw0 = rres ? rv[rb0 + lane].w : 0u, rp0 = rres ? rv[rb0 + lane].p : 0u;
        const uint32_t sp0 = sres ? stores[st0 + lane].pc : 0u, ss0 = sres ? stores[st0 + lane].slots : 0u;

        typedef typename AdjFeat<R, F>::type FeatT;
        typedef typename VgprFeat<R, (F > 0 ? F : 1)>::VF VF;
        // Row block features and targets: only the columns the program reads
        // (a wave-uniform test per column).  (Loading them one block ahead cost
        // 13 VGPRs and a wave per SIMD, and measured slower.)
        VF xf[R];
        float y[R];
        const uint32_t cols = J.cols;
        auto load_rows = [&](int64_t r0, VF (&xv)[R], float (&yv)[R]) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int64_t rw = r0 + r * 64 + lane;  // < ldx: rows are padded past nlin
                yv[r] = A.X[(int64_t)A.ycol * A.ldx + rw];
                if constexpr (F > 0) {
#pragma unroll
                    for (int f = 0; f < F; ++f) xv[r][f] = (cols >> f) & 1u ? A.X[(int64_t)f * A.ldx + rw] : 0.0f;
                }
            }
        };
        for (int64_t row0 = rbeg; row0 < rend; row0 += 64 * R) {
            load_rows(row0, xf, y);
            int row[R];
            bool valid[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                row[r] = (int)(row0 + r * 64 + lane);
                valid[r] = row[r] < rend;
            }
            FeatT ft = [&]() {
                if constexpr (F > 0) return FeatT{xf};
                else return FeatT{A.X, A.ldx, row};
            }();
            VA adj;
            for (uint32_t s = 0; s < J.nseg; ++s) {
                ConstPtr<AdjSeg> sq = sg0 + s;
                const uint32_t rb = sq->rev_begin, re = sq->rev_end, fe = sq->fwd_end;
                const uint32_t sb0 = sq->st_begin, nst = sq->st_end - sb0;
                // The segment's stores across the lanes (one window: nst <= B < 64).
                uint32_t st_pc = sp0, st_sl = ss0, ks = sb0 - st0;
                if (!sres) {
                    st_pc = stores[sb0 + lane].pc;
                    st_sl = stores[sb0 + lane].slots;
                    ks = 0;
                }
                float acc[R];
                Stk stk;
                AC::forward(code, c0, fe, fres, fwc, fwa, st_pc, st_sl, ks, ks + nst, ft, acc, stk, ps);
                if (s == 0) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        float seed = 0.0f;
                        if (valid[r]) {
                            const float d = acc[r] - y[r];
                            const float e = A.error_fn == PPI_EF_SQUARED ? d * d : fabsf(d);
                            if (!finite_f(e)) bad = true;
                            else part += e;
                            seed = A.error_fn == PPI_EF_SQUARED ? 2.0f * d : (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f));
                        }
                        adj.set(r, 0, seed);
                    }
                }
                // Reverse instructions [rb, re) through a window like the forward tape:
                // the unit's resident window, or 64-entry blocks streamed from rb.
                uint32_t blk, ww, wp;
                if (rres) {
                    blk = rb0;
                    ww = rw0;
                    wp = rp0;
                } else {
                    blk = rb & ~63u;
                    ww = rv[blk + lane].w;
                    wp = rv[blk + lane].p;
                }
                uint32_t k = rb - blk;
                for (; blk < re; blk += 64) {
                    uint32_t nw = 0, np = 0;
                    if (!rres) {
                        nw = rv[blk + 64 + lane].w;
                        np = rv[blk + 64 + lane].p;
                    }
                    const uint32_t kend = min(64u, re - blk);
                    for (; k < kend; ++k) {
                        const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)ww, (int)k);
                        const uint32_t pb = (uint32_t)__builtin_amdgcn_readlane((int)wp, (int)k);
                        const uint32_t src = (w >> 4) & 15u, dst = (w >> 8) & 31u;
                        // A zero adjoint stays zero whatever the partial (a NaN or inf
                        // partial inside a subtree that MAX/MIN/IF_THEN_ELSE did not
                        // select does not reach the constants): forward-mode semantics.
                        // The stored partial is read whatever the adjoint (in bounds),
                        // so the zero test is a select, not a branch around the read.
                        float v[R];
                        if (w & kRevStored) {
                            const uint32_t sl = (w >> 16) & (uint32_t)(B - 1);
#pragma unroll
                            for (int r = 0; r < R; ++r) {
                                const float p = ps.get(sl, r);
                                const float ad = adj.get(r, src);
                                v[r] = ad == 0.0f ? 0.0f : ad * p;
                            }
                        } else {
                            const float pc = __uint_as_float(pb);
#pragma unroll
                            for (int r = 0; r < R; ++r) {
                                const float ad = adj.get(r, src);
                                v[r] = ad == 0.0f ? 0.0f : ad * pc;
                            }
                        }
                        if ((w & 3u) == kRevAdj) {
#pragma unroll
                            for (int r = 0; r < R; ++r) adj.set(r, dst, v[r]);
                        } else {
                            float t = 0.0f;
#pragma unroll
                            for (int r = 0; r < R; ++r) t += valid[r] ? v[r] : 0.0f;
                            g[dst] += t;
                        }
                    }
                    k = 0;
                    ww = nw;
                    wp = np;
                }
            }
        }
    }
    for (uint32_t k = 0; k < J.nconst; ++k) {
        const float v = wave_sum(g[k]);
        if (lane == 0) A.gpart[(int64_t)rng * A.nconst + J.const_base + k] = v;
    }
    if (J.flags & kAdjFitness) {
        part = wave_sum(part);
        const bool any_bad = __any(bad);
        if (lane == 0) {
            A.fpart[(int64_t)rng * A.nprog + J.prog] = part;
            A.fbad[(int64_t)rng * A.nprog + J.prog] = any_bad ? 1u : 0u;
        }
    }
}

// Persistent form: the grid holds as many workgroups as are resident at once and
// every wave takes units from the queue counter (zeroed before the launch) until
// the njobs * nranges units are used up, so a wave that finishes early does not
// idle until the rest of its workgroup is done, and the launch's tail is one
// unit long.  Units are taken in order: heaviest job first within each range.
// Every wave leaves when the counter passes the end; nothing synchronises the
// block after the table staging.
template <int R, int F, int S, int AS, int PS, bool P, bool N>
__global__ void __launch_bounds__(kBlock, 3) adjoint_kernel(AdjArgs A) {
    constexpr int B = StoreCfg<PS, R>::B;
    constexpr int LDS_WAVE = PS == 1 ? B * R * 64 : 1;
    __shared__ float lds[kWaves][LDS_WAVE];
    stage_math<N>();
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = wave_index();
    const uint64_t total = (uint64_t)A.njobs * (uint64_t)A.nranges;
    for (;;) {
        // Every lane takes part in the atomic (lane 0 adds 1, the others 0) and lane
        // 0's old value is broadcast.  No branch on the lane here: with an
        // `if (lane == 0)` around the atomic, the optimiser threaded the epilogue's
        // own lane-0 test into it and sent lanes 1-63 to the broadcast without the
        // atomic, so they re-ran unit 0 without end.
        const uint32_t t = atomicAdd(A.queue, lane == 0 ? 1u : 0u);
        const uint64_t u = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
        if (u >= total) break;
        adjoint_unit<R, F, S, AS, PS, P, N>(A, (int64_t)u, &lds[wave][lane], lane);
    }
}

#ifdef PPI_ADJOINT_MISC
// Fitness with the sequential path's semantics (sequential.cc:170-189), ranges summed in order.
__global__ void adj_fit_kernel(AdjFinal F) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= F.nprog) return;
    float total = 0.0f;
    uint32_t b = 0;
    for (int r = 0; r < F.nranges; ++r) {
        total += F.fpart[(int64_t)r * F.nprog + p];
        b |= F.fbad[(int64_t)r * F.nprog + p];
    }
    const ProgMeta m = F.meta[p];
    const float fn = (float)F.nlin;
    const float as = F.alpha * (float)m.size;
    float fit;
    if (m.flags & kEmpty) {
        fit = FLT_MAX;
        b = 1;
    } else if (b) {
        fit = FLT_MAX / fn + as;
    } else if (!finite_f(total)) {
        fit = FLT_MAX;
        b = 1;
    } else {
        fit = total / fn + as;
    }
    F.fit[p] = fit;
    F.bad[p] = b;
}

__global__ void adj_grad_kernel(AdjFinal F) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= F.nconst) return;
    const AdjConst c = F.consts[k];
    float s = 0.0f;
    for (int r = 0; r < F.nranges; ++r) s += F.gpart[(int64_t)r * F.nconst + k];
    const float v = F.bad[c.prog] ? 0.0f : s / (float)F.nlin;
    F.graw[k] = v;
    if (F.gout) F.gout[(int64_t)c.prog * F.stride + c.pos] = F.gscale ? v * F.gscale[c.prog] : v;
}

// One step of normalised gradient descent with backtracking, per program: an
// evaluation that improves the best fitness is accepted (step length x 1.5),
// otherwise the step from the best point is halved.  The next point is
// cbest - lam * gbest / |gbest|; it is written into the tape for the next pass.
__global__ void adj_descend_kernel(AdjDescend D) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.nprog) return;
    const uint32_t k0 = D.prog_consts[p], k1 = D.prog_consts[p + 1];
    if (k0 == k1) return;
    const float f = D.fit[p];
    const bool b = D.bad[p] != 0;
    if (D.iter == 0) {
        D.lam[p] = D.step0;
        D.fbest[p] = f;
        for (uint32_t k = k0; k < k1; ++k) {
            D.cbest[k] = D.cur[k];
            D.gbest[k] = b ? 0.0f : D.graw[k];
        }
    } else if (!b && f < D.fbest[p]) {
        D.fbest[p] = f;
        D.lam[p] *= 1.5f;
        for (uint32_t k = k0; k < k1; ++k) {
            D.cbest[k] = D.cur[k];
            D.gbest[k] = D.graw[k];
        }
    } else {
        D.lam[p] *= 0.5f;
    }
    float nrm = 0.0f;
    for (uint32_t k = k0; k < k1; ++k) nrm += D.gbest[k] * D.gbest[k];
    nrm = sqrtf(nrm);
    const bool move = !D.last && nrm > 0.0f && finite_f(nrm);
    const float sc = move ? D.lam[p] / nrm : 0.0f;
    for (uint32_t k = k0; k < k1; ++k) {
        const float c = move ? D.cbest[k] - sc * D.gbest[k] : D.cbest[k];
        D.cur[k] = c;
        const uint32_t ins = D.consts[k].ins;
        if (ins != 0xffffffffu) D.code[ins].arg = __float_as_uint(c);
    }
}
#endif  // PPI_ADJOINT_MISC

}  // namespace dev

namespace launch {

template <int R, int F, int S, int AS, int PS>
hipError_t launch_adjoint_t(const AdjArgs& A, int grid, bool P, bool N, hipStream_t st) {
    if (P && N) hipLaunchKernelGGL((dev::adjoint_kernel<R, F, S, AS, PS, true, true>), dim3(grid), dim3(dev::kBlock), 0, st, A);
    else if (P) hipLaunchKernelGGL((dev::adjoint_kernel<R, F, S, AS, PS, true, false>), dim3(grid), dim3(dev::kBlock), 0, st, A);
    else if (N) hipLaunchKernelGGL((dev::adjoint_kernel<R, F, S, AS, PS, false, true>), dim3(grid), dim3(dev::kBlock), 0, st, A);
    else hipLaunchKernelGGL((dev::adjoint_kernel<R, F, S, AS, PS, false, false>), dim3(grid), dim3(dev::kBlock), 0, st, A);
    return hipGetLastError();
}

template <int R, int F, int S, int AS, int PS>
int adjoint_occupancy_t(bool P, bool N) {
    int n = 0;
    hipError_t e;
    if (P && N) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::adjoint_kernel<R, F, S, AS, PS, true, true>, dev::kBlock, 0);
    else if (P) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::adjoint_kernel<R, F, S, AS, PS, true, false>, dev::kBlock, 0);
    else if (N) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::adjoint_kernel<R, F, S, AS, PS, false, true>, dev::kBlock, 0);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::adjoint_kernel<R, F, S, AS, PS, false, false>, dev::kBlock, 0);
    return e == hipSuccess ? n : 0;
}

}  // namespace launch
}  // namespace ppi

#endif
