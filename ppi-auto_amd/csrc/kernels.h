// kernels.h — launch interface between the engine runtime and the HIP kernels.
#ifndef PPI_KERNELS_H
#define PPI_KERNELS_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ef.h"
#include "tape.h"

namespace ppi {

constexpr int kMaxChunk = 1024;  // programs per FP workgroup
constexpr int kDefaultMaxChunk = 400;  // the default plan's cap on programs per FP workgroup
constexpr int kPdpMaxLanes = 1024;  // PDP / DP work-group width at most (the reference's -cl-mls)

// Rows of one FP-form workgroup tile (64 lanes x R rows per lane, shared by its waves).
inline constexpr int fp_rows_per_block(int R) { return 64 * R; }

// Entries appended after the last program so that the FP form's instruction
// window (two 64-entry blocks ahead of the current instruction) stays in bounds.
constexpr int kTapePad = 192;

// Everything a kernel needs, passed by value.
struct EvalArgs {
    const ProgMeta* meta;
    const Ins* code;
    const float* X;       // column-major dataset, column c at X + c*ldx, rows padded
    int64_t ldx;
    int32_t nlin;
    int32_t ncol;
    int32_t ycol;         // ncol - 1
    int32_t nprog;
    int32_t chunk;        // programs per FP tile
    int32_t strategy;     // ppi_strategy: result semantics
    int32_t error_fn;
    int32_t reducemax;
    const float* cost;
    int32_t cost_n;
    int32_t pdp_lanes;    // reference evaluate_pdp work-group size (bad-lane bookkeeping)
    int32_t dp_groups;    // reference evaluate_dp work-group count (empty-program rule)
    int32_t dp_local;     // reference evaluate_dp work-group size (rows per group)
    float alpha;
    const int32_t* order; // PP form: lane -> program (length-sorted), may be null
    float* partial;       // FP form: [row_block][nprog]
    uint32_t* bad_lo;     // per program: min bad-row key (0xffffffff = none)
    uint32_t* bad_hi;     // per program: max bad-row key
    float* raw;           // per program: kernel-side error (before alpha*size for PP/PDP)
    float* vec;           // per program: reported error
    // Ordered form: the rows run in passes of nrb_pass row blocks starting at row block rb0
    // (rb0 also offsets the FP form's row blocks; 0 outside the ordered form).
    int32_t rb0, nrb_pass;
    float* rowerr;        // [row block - rb0][nprog][64*R]: per-row errors of one pass
    float* osum;          // [nprog] running sum (REDUCEMAX: maximum) in row order
    uint32_t* ostop;      // [nprog] 1 once a row's error was not finite
    const EfIns* ef;      // PPI_EF_EXPR: the compiled ERROR(X,Y) (ef.h), ef_n operations
    int32_t ef_n;
    // The row-tile partials the recomputation may start from: part_nrb tiles of part_rows rows
    // in `partial` (FP and batched forms); 0 where the form keeps none.
    int32_t part_nrb, part_rows;
    // Deep list of the recomputation (deep_cap > 0): programs whose partition-exact result needs
    // every row are queued here, their rows computed across the device and then reduced in the
    // reference's partition, instead of by one workgroup each.
    uint32_t* deep_count;  // entries queued (may exceed deep_cap: the rest stay in the recheck)
    int32_t* deep_prog;    // [deep_cap] program id (the engine's index, or the batch's global one)
    int32_t deep_cap;
};

// Batched form: one (population, dataset) pair as the device sees it.
struct PairDev {
    const float* X;        // the pair's column-major dataset
    int64_t ldx;
    int32_t nlin, ncol;
    int32_t prog_base;     // first program of the pair in the concatenated population
    int32_t nprog;
    int64_t partial_base;  // first partial of the pair: [row_block][program] block of nrb*nprog
    int32_t nrb;           // row tiles of rows_per_tile = 64*R rows (R: the pair's rows per lane)
    int32_t pdp_lanes, dp_groups;
    int32_t dp_local;
    int32_t rows_per_tile;
};

// One work item: a row tile of a pair x a program range [p0, p1) of that pair.
struct WorkItem {
    int32_t pair;
    int32_t row_tile;
    int32_t p0, p1;
};

struct BatchArgs {
    EvalArgs base;          // tape, meta (all pairs), result and work buffers; per-pair fields overwritten
    const PairDev* pairs;
    const WorkItem* items;
    int32_t n_items;
    uint32_t* counter;      // work-queue head, reset to 0 before every launch
};

// The deep list's row errors: entry k's rows at err + k * stride.  pairs / pair_of: the batched
// form's pair table (null for the engine).
struct DeepArgs {
    EvalArgs base;
    const PairDev* pairs;
    const int32_t* pair_of;
    float* err;
    int64_t stride;     // >= the longest dataset's rows
    int32_t max_rows;   // rows of the longest dataset
};

// Reverse-mode adjoint: waves take (job, row range) units from a queue counter.
struct AdjArgs {
    const AdjJob* jobs;
    const AdjSeg* segs;
    const RevIns* rev;
    const AdjStore* stores;
    const Ins* code;
    const int32_t* order;  // unit u -> job order[u % njobs] (heaviest first), range u / njobs
    int32_t njobs;
    int32_t nranges;
    int64_t range_rows;    // rows per range, a multiple of 64*R
    const float* X;        // column-major dataset, rows padded to ldx
    int64_t ldx;
    int32_t nlin, ncol, ycol;
    int32_t error_fn;      // PPI_EF_ABS or PPI_EF_SQUARED
    int32_t nprog;
    int64_t nconst;
    float* gpart;          // [range][nconst]: sum over the range's rows of dEF/dpred * dpred/dconst
    float* fpart;          // [range][nprog]: sum of finite errors
    uint32_t* fbad;        // [range][nprog]: 1 if a row's error is not finite
    uint32_t* queue;       // next unit (zeroed before each launch); njobs * nranges < 2^32 - grid waves
};

// Per-program and per-constant epilogue of the adjoint.
struct AdjFinal {
    const ProgMeta* meta;
    const AdjConst* consts;
    int32_t nprog;
    int64_t nconst;
    int32_t nranges;
    int32_t nlin;
    float alpha;
    const float* fpart;
    const uint32_t* fbad;
    const float* gpart;
    float* fit;            // [nprog] fitness, sequential-path semantics (sequential.cc:170-189)
    uint32_t* bad;         // [nprog] 1: empty, or an error is not finite -> gradient 0
    float* graw;           // [nconst] d fitness / d constant
    const float* gscale;   // [nprog] upstream gradient, null = 1
    float* gout;           // [nprog * stride] ephemeral layout, null = not written
    int32_t stride;
};

// On-device constant refinement: one update per program and call.
struct AdjDescend {
    const uint32_t* prog_consts;  // nprog + 1 offsets into the constant table
    const AdjConst* consts;
    int32_t nprog;
    const float* fit;
    const uint32_t* bad;
    const float* graw;
    float* cur;    // [nconst] constants being evaluated
    float* cbest;  // [nconst] best constants so far
    float* gbest;  // [nconst] gradient at cbest
    float* fbest;  // [nprog]
    float* lam;    // [nprog] step length
    Ins* code;     // forward tape: the constants' instruction arguments are rewritten
    int32_t iter;
    float step0;
    int32_t last;  // 1: write back cbest and stop
};

// Adjoint kernel shapes: R rows per lane, F features in VGPRs (0: from the dataset),
// S stack slots, AS adjoint slots, PS partial store (0: VGPRs, 1: LDS).
bool adjoint_shape_supported(int R, int F, int S, int AS, int PS);
int adjoint_store_partials(int PS, int R);
hipError_t launch_adjoint(const AdjArgs& A, int grid, int R, int F, int S, int AS, int PS, bool P, bool N,
                          hipStream_t st);
int adjoint_occupancy(int R, int F, int S, int AS, int PS, bool P, bool N);
// The two-rows-per-lane shapes (their own instantiation unit).
bool adjoint_shape_supported_r2(int F, int S, int AS, int PS);
hipError_t launch_adjoint_r2(const AdjArgs& A, int grid, int F, int S, int AS, int PS, bool P, bool N, hipStream_t st);
int adjoint_occupancy_r2(int F, int S, int AS, int PS, bool P, bool N);
hipError_t launch_adjoint_finalize(const AdjFinal& F, hipStream_t st);
hipError_t launch_adjoint_descend(const AdjDescend& D, hipStream_t st);

bool shape_supported(int R, int F, int S);
bool batch_shape_supported(int R, int F, int S);
hipError_t launch_batch(const BatchArgs& B, int grid, int R, int F, int S, bool P, bool N, hipStream_t st);
int batch_occupancy(int R, int F, int S, bool P, bool N);
hipError_t launch_batch_finalize(const BatchArgs& B, const int32_t* pair_of, int32_t n_total, hipStream_t st);
bool pdp_shape_supported(int R, int F, int S);
// rows: the ordered form's FP pass (per-row errors to A.rowerr instead of partials).
hipError_t launch_fp(const EvalArgs& A, int R, int F, int S, bool P, bool N, hipStream_t st, bool rows = false);
// Ordered form: continues every program's row-order sum over the pass's rows; last: apply the epilogue.
hipError_t launch_ordered_sum(const EvalArgs& A, int rows_per_block, bool last, hipStream_t st);
hipError_t launch_pdp(const EvalArgs& A, int R, int F, int S, bool P, bool N, hipStream_t st);
hipError_t launch_pp(const EvalArgs& A, bool P, bool N, hipStream_t st);
hipError_t launch_finalize(const EvalArgs& A, int nrb, hipStream_t st);
// Recomputes, in the reference's own partition and order (PDP lanes, DP work-groups, row order), the
// programs whose epilogue marked them (raw[p] holding the recheck NaN); unmarked programs cost one load.
hipError_t launch_recheck(const EvalArgs& A, bool P, bool N, hipStream_t st);
hipError_t launch_batch_recheck(const BatchArgs& B, const int32_t* pair_of, int32_t n_total, bool P, bool N,
                                hipStream_t st);
// The recheck's deep list (D.base.deep_cap > 0), after launch_recheck / launch_batch_recheck: the
// queued programs' row errors over `grid` workgroups, then one workgroup per entry for the
// partition-exact result.  Programs are queued only from datasets of at least kDeepMinRows rows.
constexpr int32_t kDeepCap = 16;
constexpr int32_t kDeepMinRows = 8192;
hipError_t launch_deep(const DeepArgs& D, int grid, bool P, bool N, hipStream_t st);
hipError_t launch_predict(const EvalArgs& A, int F, int S, bool P, bool N, float* out, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st);
// PPI_EF_EXPR: empty programs get their result, every other one is marked for the exact-order
// recomputation (launch_recheck), which evaluates the expression per row.
hipError_t launch_ef_prepare(const EvalArgs& A, hipStream_t st);
// Folded constant subtrees: out[j] = value of fold program j (meta[j], code).
hipError_t launch_fold(const ProgMeta* meta, const Ins* code, int32_t n, float* out, bool P, bool N, hipStream_t st);
// Derived columns first .. first+count-1 of X: column first+j = unary(ops[j], column srcs[j]).
hipError_t launch_derive(float* X, int64_t ldx, const uint32_t* ops, const int32_t* srcs, int32_t first,
                         int32_t count, bool P, bool N, hipStream_t st);

}  // namespace ppi

#endif
