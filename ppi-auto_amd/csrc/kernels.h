// kernels.h — launch interface between the engine runtime and the HIP kernels.
#ifndef PPI_KERNELS_H
#define PPI_KERNELS_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tape.h"

namespace ppi {

constexpr int kMaxChunk = 1024;  // programs per FP workgroup

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
    float alpha;
    const int32_t* order; // PP form: lane -> program (length-sorted), may be null
    float* partial;       // FP form: [row_block][nprog]
    uint32_t* bad_lo;     // per program: min bad-row key (0xffffffff = none)
    uint32_t* bad_hi;     // per program: max bad-row key
    float* raw;           // per program: kernel-side error (before alpha*size for PP/PDP)
    float* vec;           // per program: reported error
};

// Batched form: one (population, dataset) pair as the device sees it.
struct PairDev {
    const float* X;        // the pair's column-major dataset
    int64_t ldx;
    int32_t nlin, ncol;
    int32_t prog_base;     // first program of the pair in the concatenated population
    int32_t nprog;
    int64_t partial_base;  // first partial of the pair: [row_block][program] block of nrb*nprog
    int32_t nrb;           // row tiles of 64*R rows
    int32_t pdp_lanes, dp_groups;
    int32_t pad_;
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

bool shape_supported(int R, int F, int S);
bool batch_shape_supported(int R, int F, int S);
hipError_t launch_batch(const BatchArgs& B, int grid, int R, int F, int S, bool P, bool N, hipStream_t st);
int batch_occupancy(int R, int F, int S, bool P, bool N);
hipError_t launch_batch_finalize(const BatchArgs& B, const int32_t* pair_of, int32_t n_total, hipStream_t st);
bool pdp_shape_supported(int R, int F, int S);
hipError_t launch_fp(const EvalArgs& A, int R, int F, int S, bool P, bool N, hipStream_t st);
hipError_t launch_pdp(const EvalArgs& A, int R, int F, int S, bool P, bool N, hipStream_t st);
hipError_t launch_pp(const EvalArgs& A, bool P, bool N, hipStream_t st);
hipError_t launch_finalize(const EvalArgs& A, int nrb, hipStream_t st);
hipError_t launch_predict(const EvalArgs& A, int F, int S, bool P, bool N, float* out, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st);

}  // namespace ppi

#endif
