/*
 * ppi_engine.h — C ABI of the MI355X (gfx950) program-evaluation engine for PPI.
 *
 * The engine evaluates a population of grammar-generated prefix programs over a
 * float32 dataset and returns one error per program, which is the hot path of
 * PPI's src/interpreter (reference: src/interpreter/accelerator.{h,cc,cl} and
 * src/interpreter/sequential.{h,cc}).  The C++ drop-in shim (acc_interpret_*,
 * seq_interpret_*) is a thin layer over these entry points; see
 * include/ppi_interpreter.h.
 *
 * Conventions
 *   - Plain C types only; no exceptions cross the ABI.  Every entry point
 *     returns a ppi_status; ppi_last_error() gives a message for the calling
 *     thread.
 *   - Host arrays are owned by the caller.  The engine copies what it needs.
 *   - One engine object holds one dataset and one resident program set on one
 *     device; it is not re-entrant (the reference keeps one static interpreter
 *     per process, accelerator.cc:52).
 */
#ifndef PPI_ENGINE_H
#define PPI_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PPI_ABI_VERSION 1

typedef enum ppi_status {
    PPI_OK = 0,
    PPI_E_ARG = 1,         /* invalid argument */
    PPI_E_STATE = 2,       /* call out of order (no dataset, no programs, ...) */
    PPI_E_PROGRAM = 3,     /* a program cannot be compiled (too many stack slots, ...) */
    PPI_E_DEVICE = 4,      /* HIP runtime error */
    PPI_E_NOMEM = 5,       /* host or device allocation failed */
    PPI_E_UNSUPPORTED = 6, /* feature not available in this build */
    PPI_E_NODEVICE = 7     /* no HIP device present */
} ppi_status;

/*
 * Engine opcodes.  A grammar's Symbol values are mapped onto these by a symbol
 * table (ppi_symbol).  Operator semantics follow src/interpreter/interpreter_core
 * and src/interpreter/functions.h of the reference; operand order follows the
 * interpreter's stack: for a binary operator the FIRST child in prefix order is
 * the left operand (interpreter_core:48-51, stack[top] is the left operand).
 */
typedef enum ppi_op {
    PPI_OP_UNKNOWN = 0,   /* unrecognised symbol: pushes NaN (sequential.cc:160-162) */
    PPI_OP_CONST = 1,     /* T_CONST: pushes ephemeral[i] (sequential.cc:155-158) */
    PPI_OP_ATTRIBUTE = 2, /* T_ATTRIBUTE: pushes input[row][(int)ephemeral[i]] (sequential.cc:150-153) */
    PPI_OP_VALUE = 3,     /* named or integer constant (T_PI, T_0, T_N7, ...): pushes ppi_symbol.value */

    /* ternary */
    PPI_OP_IF_THEN_ELSE = 10,
    /* binary */
    PPI_OP_AND = 11, PPI_OP_OR = 12, PPI_OP_XOR = 13,
    PPI_OP_GREATER = 14, PPI_OP_GREATEREQUAL = 15, PPI_OP_LESS = 16, PPI_OP_LESSEQUAL = 17,
    PPI_OP_EQUAL = 18, PPI_OP_NOTEQUAL = 19,
    PPI_OP_ADD = 20, PPI_OP_SUB = 21, PPI_OP_MULT = 22, PPI_OP_MEAN = 23,
    PPI_OP_MAX = 24, PPI_OP_MIN = 25, PPI_OP_MOD = 26, PPI_OP_POW = 27, PPI_OP_DIV = 28,
    /* unary */
    PPI_OP_NOT = 40, PPI_OP_ABS = 41, PPI_OP_POW2 = 42, PPI_OP_POW3 = 43, PPI_OP_POW4 = 44,
    PPI_OP_POW5 = 45, PPI_OP_NEG = 46, PPI_OP_ROUND = 47, PPI_OP_CEIL = 48, PPI_OP_FLOOR = 49,
    PPI_OP_EXP = 50, PPI_OP_EXP10 = 51, PPI_OP_EXP2 = 52, PPI_OP_SIN = 53, PPI_OP_COS = 54,
    PPI_OP_STEP = 55, PPI_OP_SIGN = 56, PPI_OP_LOGISTIC = 57, PPI_OP_GAMMA = 58, PPI_OP_GAUSS = 59,
    PPI_OP_SQRT = 60, PPI_OP_LOG = 61, PPI_OP_LOG10 = 62, PPI_OP_LOG2 = 63, PPI_OP_TAN = 64,

    PPI_OP_COUNT_ = 65
} ppi_op;

/* One symbol-table entry: grammar Symbol value -> engine opcode (+ value for PPI_OP_VALUE). */
typedef struct ppi_symbol {
    int32_t symbol;
    int32_t op;
    float value;
} ppi_symbol;

/* Per-row error function ERROR(X,Y) (definitions.h:18, CMakeLists.txt:19-33). */
typedef enum ppi_error_fn {
    PPI_EF_ABS = 0,        /* fabs((X)-(Y))                       (reference default) */
    PPI_EF_SQUARED = 1,    /* ((X)-(Y))*((X)-(Y)) */
    PPI_EF_NEQ = 2,        /* ((X)!=(Y))                          (problem/iris/README.md) */
    PPI_EF_COST = 3,       /* COST[(int)X][(int)Y]                (default when COST is given) */
    PPI_EF_COST_CLAMP = 4, /* COST[(int)fminf(fmaxf(X,0),N-1)][(int)Y] */
    PPI_EF_EXPR = 5        /* any other ERROR(X,Y): ppi_config.ef_expr, a C expression in X and Y compiled at
                              configure (the types, operators and functions ppi_ef_evaluate documents); every
                              program then takes the exact-order path (the reference's own partition and row
                              order for the strategy), with the expression interpreted per row */
} ppi_error_fn;

/* Result semantics: which reference path's error vector and ranking to reproduce. */
typedef enum ppi_strategy {
    PPI_STRATEGY_SEQ = 0,  /* seq_interpret (sequential.cc:117-209) */
    PPI_STRATEGY_PP = 1,   /* evaluate_pp + host epilogue (accelerator.cl:7-86, accelerator.cc:843-889) */
    PPI_STRATEGY_FP = 2,   /* evaluate_dp ("DP", fitness-parallel) + host reduction (accelerator.cl:88-171, accelerator.cc:756-840) */
    PPI_STRATEGY_PDP = 3   /* evaluate_pdp (accelerator.cl:173-268) */
} ppi_strategy;

/* Kernel form used to compute the result (any form can serve any strategy's semantics). */
typedef enum ppi_form {
    PPI_FORM_AUTO = 0,  /* autotuned */
    PPI_FORM_PP = 1,    /* one lane per program, rows sequential */
    PPI_FORM_FP = 2,    /* row tiles x program chunks, rows across lanes, programs wave-uniform */
    PPI_FORM_PDP = 3,   /* one workgroup per program, lanes over rows */
    PPI_FORM_ORDERED = 4 /* FP row tiles write per-row errors; each program's errors are then summed in row
                            order, so sums are bit-identical to the sequential loop (SEQ and PP semantics only) */
} ppi_form;

typedef struct ppi_config {
    int32_t strategy;       /* ppi_strategy */
    int32_t form;           /* ppi_form */
    int32_t error_fn;       /* ppi_error_fn */
    int32_t reducemax;      /* REDUCEMAX: max(EF*nlin) instead of the sum (definitions.h:21) */
    int32_t protected_ops;  /* PROTECTED (functions.h:99-193) */
    int32_t native_math;    /* NATIVE: fast float transcendentals instead of the sequential path's
                               double-rounded ones (functions.h:19-57) */
    int32_t cost_n;         /* COST matrix is cost_n x cost_n, row-major in cost[] */
    const float* cost;
    /* Reference launch geometry that shapes result semantics (accelerator.cc:272-355).
       <= 0 selects what acc_interpret_init computes on this device. */
    int32_t max_local_size; /* -cl-mls: 0 = the device's limit; at most 1024 (PPI_E_ARG above) */
    int32_t compute_units;  /* CL_DEVICE_MAX_COMPUTE_UNITS */
    /* Tuning; 0 = autotune. */
    int32_t rows_per_lane;  /* 1, 2, 4 or 8 */
    int32_t programs_per_tile;
    int32_t ordered_pass_rows; /* PPI_FORM_ORDERED: rows per pass of stored row errors (0: min(16 GiB, half the
                                  free device memory) of errors) */
    int32_t no_constant_folding; /* 1: evaluate constant subtrees on every row instead of once per upload
                                    (results are the same bits either way) */
    const char* ef_expr;    /* PPI_EF_EXPR: the text of ERROR(X,Y) (copied at configure) */
} ppi_config;

typedef struct ppi_stats {
    double upload_ms;       /* host pack + H2D of the last program set */
    double kernel_ms;       /* device time of the last evaluation (all kernels, HIP events) */
    double main_kernel_ms;  /* device time of the evaluation kernel alone */
    double download_ms;     /* D2H of the results */
    double node_evals;      /* sum(size) * nlin of the last evaluation (ppi.cc:588,652) */
    double executed_nodes;  /* tape instructions actually executed * rows */
    int32_t form;           /* form that ran */
    int32_t rows_per_lane;
    int32_t programs_per_tile;
    int32_t max_slots;      /* register stack slots used by the deepest program */
    int32_t derived_columns;/* unary-of-input columns precomputed for the FP form (this dataset) */
    int32_t folded_subtrees;/* constant subtrees of the last program set evaluated once at upload */
    int32_t folded_nodes;   /* operator nodes inside them */
    int32_t derived_refills;/* times the full derived-column table was refilled for a population it no longer
                               covered, since the dataset or configuration was set */
} ppi_stats;

typedef struct ppi_engine ppi_engine;

/* Library / device management. */
int32_t ppi_abi_version(void);
const char* ppi_last_error(void);
ppi_status ppi_device_count(int32_t* count);

/* acc_interpret_init equivalent (accelerator.cc:557-648): create, configure, stage dataset. */
ppi_status ppi_engine_create(ppi_engine** out, int32_t device);
void ppi_engine_destroy(ppi_engine* eng);
ppi_status ppi_engine_configure(ppi_engine* eng, const ppi_config* cfg,
                                const ppi_symbol* table, int32_t ntable);
/* Dataset as the reference receives it: float** input, input[row][col], y = column ncol-1
   (main.cc:32-122, accelerator.cc:363-430).  Copied once. */
ppi_status ppi_engine_set_dataset(ppi_engine* eng, const float* const* input, int32_t nlin, int32_t ncol);
/* Same, from one row-major block of nlin*ncol floats. */
ppi_status ppi_engine_set_dataset_rowmajor(ppi_engine* eng, const float* data, int32_t nlin, int32_t ncol);

/* Program upload (the H2D half of acc_interpret, accelerator.cc:668-684): phenotype,
   ephemeral and size exactly as ppi.cc:312-314 lays them out (program i occupies
   [i*stride, i*stride + size[i])).  Programs are compiled into a packed device tape. */
ppi_status ppi_engine_upload(ppi_engine* eng, const int32_t* phenotype, const float* ephemeral,
                             const int32_t* size, int32_t stride, int32_t n_programs);

/* Launch the evaluation of the resident programs (asynchronous; accelerator.cc:694-733). */
ppi_status ppi_engine_launch(ppi_engine* eng);
/* Wait for the launch and fetch results (accelerator.cc:743-889):
     vector[i] = the strategy's error of program i including alpha*size[i];
     if index != NULL and best_size != NULL: index[0..*best_size) = best programs by the
     strategy's ranking rule (*best_size may be lowered as the reference does). */
ppi_status ppi_engine_wait(ppi_engine* eng, float alpha, float* vector,
                           int32_t* index, int32_t* best_size);
/* upload + launch + wait. */
ppi_status ppi_engine_evaluate(ppi_engine* eng, const int32_t* phenotype, const float* ephemeral,
                               const int32_t* size, int32_t stride, int32_t n_programs, float alpha,
                               float* vector, int32_t* index, int32_t* best_size);
/* Prediction mode (ppp -pred; accelerator.cl:72-75): out[row] = value of program 0 on every row. */
ppi_status ppi_engine_predict(ppi_engine* eng, const int32_t* phenotype, const float* ephemeral,
                              int32_t size, float* out);
/* ppi_engine_evaluate without allocating: no counterpart in the reference.  For callers that
   cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc).  Same arguments and results as ppi_engine_evaluate (vector and, when
   index/best_size are given, the ranking), for the FP and PDP forms (PPI_E_UNSUPPORTED for the
   others).  The programs, constant-subtree programs, partials, bad-row records and results live
   in a static 64 MiB device array per GPU (held until the process ends); no device or pinned host
   memory is allocated (the host compile uses ordinary heap memory) and no derived columns are
   added.  A population whose buffers exceed 64 MiB (about 16 M programs x row tiles of 256 rows,
   e.g. 30,000 programs x 100,000 rows fits) is rejected with PPI_E_ARG.  The call is
   synchronous; calls that overlap on one GPU share the array and race.  The engine's resident
   program set (upload / launch / wait) is left as it was. */
ppi_status ppi_engine_evaluate_noalloc(ppi_engine* eng, const int32_t* phenotype, const float* ephemeral,
                                      const int32_t* size, int32_t stride, int32_t n_programs, float alpha,
                                      float* vector, int32_t* index, int32_t* best_size);
ppi_status ppi_engine_get_stats(const ppi_engine* eng, ppi_stats* out);

/*
 * User error function ERROR(X,Y) (definitions.h:18; -DEF=<expression>, CMakeLists.txt:26-33),
 * evaluated on the host as the device evaluates it: out[i] = ERROR(X[i], Y[i]).  The
 * expression is C over float X and Y with C's typing (int, float and double literals and the
 * usual arithmetic conversions; + - * / %, comparisons, && || !, ?:, casts to int, float and
 * double) and the functions sequential.cc sees through <cmath>: fabs sqrt exp log pow sin cos
 * tan floor ceil round trunc fmin fmax (double), fabsf sqrtf floorf ceilf roundf truncf fminf
 * fmaxf (float), abs (its argument's type).  PPI_E_ARG, with the reason in ppi_last_error(),
 * for anything else (the C library's float transcendentals expf ... powf among them: they are
 * not correctly rounded, so their results cannot be reproduced).  No reference counterpart (the
 * reference compiles the expression into its kernels); the drop-in uses it to check, at init,
 * that the compiled expression gives the macro's own values.
 */
ppi_status ppi_ef_evaluate(const char* text, const float* X, const float* Y, int32_t n, float* out);

/* Device-resident timing loop used by benchmarks: run `iters` evaluations of the resident
   programs back to back on the engine stream, return mean device ms per evaluation. */
ppi_status ppi_engine_time(ppi_engine* eng, int32_t iters, double* ms_per_eval);

/*
 * Batched evaluation: many (population, dataset) pairs in one persistent launch.
 * No counterpart in the reference (which evaluates one population on one dataset
 * per call); it serves GE runs that evaluate many islands or problems on one GPU.
 * Datasets are staged once; populations are copied when added and compiled by
 * build(), all pending ones at once on a pool of host threads, which then plans
 * the work (autotuning rows per lane and work-item size when
 * ppi_config.rows_per_lane is 0) and captures upload-free replay as a hipGraph.
 * A program that cannot be compiled (more than 16 register slots) fails build().
 */
typedef struct ppi_batch ppi_batch;

typedef struct ppi_batch_stats {
    double graph_ms;        /* last replay: whole graph (reset, kernel, finalize, D2H), device time */
    double kernel_ms;       /* evaluation kernels alone: direct FP launches + persistent queue (ppi_batch_time) */
    double node_evals;      /* sum over pairs of sum(size) * nlin */
    double build_ms;        /* compile + upload + plan + autotune + capture */
    int32_t n_pairs;
    int32_t n_items;
    int32_t grid;           /* persistent workgroups of the launch carrying the most work */
    int32_t rows_per_lane;  /* R of that launch (each pair's own: ppi_batch_pair_shape) */
    int32_t feature_slots;  /* F of that launch's kernel shape (0 = attributes read from global memory) */
    int32_t autotuned;      /* 1 if some pair's shape was picked by measurement */
    double tune_ms[4];      /* sampled ms per 10^9 instruction x rows, candidates R = 2, 4, 8 (0 = not
                               tried), of the tuning key carrying the most work */
    int32_t derived_columns; /* derived columns (unary operators over input columns) computed, all datasets */
    int32_t n_shapes;       /* persistent launches per replay: one per kernel shape in use */
    int32_t tuned_keys;     /* tuning keys (stack class x program length x rows) timed by this build
                               (keys decided by an earlier build of the batch keep their shape) */
    int32_t direct_pairs;   /* pairs evaluated by their own FP launch instead of the persistent queue */
} ppi_batch_stats;

ppi_status ppi_batch_create(ppi_batch** out, int32_t device);
void ppi_batch_destroy(ppi_batch* b);
ppi_status ppi_batch_configure(ppi_batch* b, const ppi_config* cfg, const ppi_symbol* table, int32_t ntable);
/* Row-major nlin x ncol dataset, target in the last column; returns its id. */
ppi_status ppi_batch_add_dataset(ppi_batch* b, const float* data, int32_t nlin, int32_t ncol, int32_t* dataset_id);
/* A population evaluated on dataset `dataset_id` (ppi.cc:312-314 layout); returns its id.  The
   arrays are copied (a size outside [0, stride] is rejected here); build() compiles them. */
ppi_status ppi_batch_add_population(ppi_batch* b, int32_t dataset_id, const int32_t* phenotype,
                                    const float* ephemeral, const int32_t* size, int32_t stride, int32_t n,
                                    int32_t* pop_id);
/* Drop all populations (datasets stay staged). */
ppi_status ppi_batch_clear_populations(ppi_batch* b);
ppi_status ppi_batch_build(ppi_batch* b);
ppi_status ppi_batch_launch(ppi_batch* b);
ppi_status ppi_batch_wait(ppi_batch* b);
/* Results of one population after wait(), with the configured strategy's semantics and ranking. */
ppi_status ppi_batch_results(ppi_batch* b, int32_t pop_id, float alpha, float* vector, int32_t* index,
                             int32_t* best_size);
/* `iters` graph replays back to back; mean device ms per replay (also times the kernel alone). */
ppi_status ppi_batch_time(ppi_batch* b, int32_t iters, double* ms_per_run);
ppi_status ppi_batch_get_stats(const ppi_batch* b, ppi_batch_stats* out);
/* The kernel shape the autotuner gave population `pop_id` after build(): rows per lane, LDS
   feature slots, register-stack slots (any pointer may be NULL).  No reference counterpart. */
ppi_status ppi_batch_pair_shape(const ppi_batch* b, int32_t pop_id, int32_t* rows_per_lane,
                                int32_t* feature_slots, int32_t* stack_slots);
/* Which pairs the next build() evaluates by their own FP launch (the engine's kernel and
   chunking) instead of the persistent queue: those with at least `min_items` work items of
   (row tile, program), row tiles of 64 x rows-per-lane rows.  0 = the default (96 x 64 items
   per compute unit: the pair's launch fills the device), a negative value = none, 1 = every
   pair.  Results are the same either way.  No reference counterpart. */
ppi_status ppi_batch_set_direct_threshold(ppi_batch* b, int64_t min_items);

/*
 * Reverse-mode derivatives of the fitness with respect to the ephemeral constants.
 * No counterpart in the reference (its T_CONST values are set by the genome,
 * ppi.cc:124-137, and never refined); it serves constant refinement on the device
 * while the grammar and the GE operators stay as they are.
 *
 * Shaped like an autograd function:
 *   forward   fitness[i] of each program (sequential-path semantics, the engine's FP
 *             form), and the programs are kept for backward;
 *   backward  grad[i*stride + j] = grad_fitness[i] * d fitness[i] / d ephemeral[i*stride + j]
 *             for every T_CONST position j (0 elsewhere), by a reverse sweep over the
 *             program tape;
 *   descend   `steps` normalised gradient steps with per-program backtracking, all on
 *             the device; returns the refined constants and the plain evaluator's
 *             fitness of the refined programs.
 * Supported error functions: PPI_EF_ABS and PPI_EF_SQUARED with the sum reduction.
 * A program whose fitness is a non-finite sentinel (some row's error is not finite,
 * or size 0) gets gradient 0.  Constants under a comparison, a logical operator,
 * ROUND/CEIL/FLOOR/STEP/SIGN or the condition of IF_THEN_ELSE have gradient 0.
 */
typedef struct ppi_adjoint ppi_adjoint;

typedef enum ppi_partials {
    PPI_PARTIALS_AUTO = 0,       /* pick by measurement */
    PPI_PARTIALS_RECOMPUTE = 1,  /* 16 / R partials per row in registers (R rows per lane); longer programs
                                    re-run the forward pass */
    PPI_PARTIALS_STORE = 2       /* 32 / R partials per row in LDS; most programs need one forward pass */
} ppi_partials;

typedef struct ppi_adjoint_stats {
    double compile_ms;       /* adjoint tape compile + upload (last backward/descend) */
    double backward_ms;      /* last backward: adjoint kernel + epilogue, device time */
    double kernel_ms;        /* adjoint kernel alone, device time (ppi_adjoint_time) */
    double descend_ms;       /* last descend: all steps, device time */
    double node_evals;       /* sum(size) * nlin */
    double forward_ins;      /* executed forward tape instructions x rows, all jobs and segments */
    double reverse_ins;      /* executed reverse instructions x rows */
    int32_t n_jobs;
    int32_t n_consts;        /* constants with a non-zero derivative path */
    int32_t max_segments;
    int32_t partials;        /* ppi_partials used last */
    int32_t rows_per_lane;
    int32_t ranges;          /* row ranges per job */
    int32_t descend_steps;
    int32_t descend_reverted; /* programs whose refined constants did not lower the plain fitness (kept) */
    int32_t feature_slots;   /* 12: a row's first 12 columns held in registers; 0: read from the dataset */
    int32_t pad_;
    double tune_ms[8];       /* sampled ms per candidate ((R - 1) * 2 + (STORE ? 1 : 0)) * 2 + (feature_slots == 0),
                                R rows per lane; 0 = not sampled, -2 = no kernel shape holds the programs so */
} ppi_adjoint_stats;

ppi_status ppi_adjoint_create(ppi_adjoint** out, int32_t device);
void ppi_adjoint_destroy(ppi_adjoint* a);
/* Same configuration as an engine (strategy, form and rows_per_lane apply to forward). */
ppi_status ppi_adjoint_configure(ppi_adjoint* a, const ppi_config* cfg, const ppi_symbol* table, int32_t ntable);
ppi_status ppi_adjoint_set_dataset_rowmajor(ppi_adjoint* a, const float* data, int32_t nlin, int32_t ncol);
ppi_status ppi_adjoint_forward(ppi_adjoint* a, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                               int32_t stride, int32_t n_programs, float alpha, float* fitness);
/* grad_fitness: n_programs upstream gradients, or NULL for all ones.  grad: n_programs*stride. */
ppi_status ppi_adjoint_backward(ppi_adjoint* a, const float* grad_fitness, int32_t partials, float* grad);
/* ephemeral_out: n_programs*stride (the forward's ephemerals with refined T_CONST values);
   fitness_out: n_programs, the plain evaluator's fitness of the returned programs; either may be
   NULL.  A program whose refined constants do not lower that fitness keeps its original ones, so
   no program comes out worse.  The kept programs are updated. */
ppi_status ppi_adjoint_descend(ppi_adjoint* a, int32_t steps, float step, int32_t partials, float* ephemeral_out,
                               float* fitness_out);
/* `iters` backward passes back to back (no host transfers); mean device ms per pass. */
ppi_status ppi_adjoint_time(ppi_adjoint* a, int32_t iters, int32_t partials, double* ms_per_pass);
ppi_status ppi_adjoint_get_stats(const ppi_adjoint* a, ppi_adjoint_stats* out);
/* Rows per lane of the adjoint kernel: 0 = chosen by measurement with the partial store (default),
   1, or 2 (programs of at most 5 stack and 5 adjoint slots; otherwise backward returns
   PPI_E_UNSUPPORTED).  Kept across forward calls. */
ppi_status ppi_adjoint_set_rows_per_lane(ppi_adjoint* a, int32_t rows_per_lane);

/*
 * Evaluation with the reference's epilogue on the device: upload, launch, then alpha*size and the
 * strategy's ranking (best_individual + PickNBest, or PickNBest over the vector: accelerator.cl:270-313,
 * accelerator.cc:843-889, sequential.cc:197-208) computed by kernels; only the vector and the index
 * come back.  Same arguments and results as ppi_engine_evaluate (best_size at most 64).
 */
ppi_status ppi_engine_evaluate_ranked(ppi_engine* eng, const int32_t* phenotype, const float* ephemeral,
                                      const int32_t* size, int32_t stride, int32_t n_programs, float alpha,
                                      float* vector, int32_t* index, int32_t* best_size);
/* The resident compiled tape (test and inspection): meta[4*i..] = {code_off, ncode, size, flags} per
   program, code[2*k..] = {code, arg} per instruction; *n_code receives the instruction count (when
   code is NULL only the count is returned). */
ppi_status ppi_engine_resident_tape(const ppi_engine* eng, uint32_t* meta, uint32_t* code, int64_t* n_code);

/*
 * Grammatical evolution on the device: the genome -> phenotype decoder (ppi.cc:106-182), the
 * generation loop of ppi_evolve (ppi.cc:823-945: elitism, tournament selection, one- and two-point
 * crossover, bit-flip and shrink mutation, ppi.cc:422-452,680-814) and the per-generation
 * evaluation and best-individual tracking of ppi_evaluate (ppi.cc:570-640), with the programs
 * compiled into the engine's tape by the device (no host round trip).  Genomes, phenotypes,
 * tapes and fitness stay in HBM; the host reads back only what it asks for.
 *
 * Random numbers reproduce the reference's: XorShift128+ (util/Random.h:130-162), one generator
 * per OpenMP thread seeded seed ^ thread (ppi.cc:388-396), each thread drawing for the
 * iterations the static schedule gives it.  With the same seed, thread count, grammar, dataset
 * and evaluation semantics, the device run follows the reference's run genome for genome (the
 * evaluation must give the reference's bits: SEQ or PP strategy with the ordered form).
 * Islands are the caller's: immigrants are handed over with ppi_ge_set_immigrants.
 */
typedef struct ppi_ge ppi_ge;

/* The decoder's grammar: the grammar header's gramatica[] and tamanhos[] (script/read_grammar.py). */
typedef struct ppi_ge_grammar {
    int32_t n_nonterminals;        /* non-terminals are Symbol values 0 .. n-1; NT_INICIAL = 0 starts */
    const int32_t* n_rules;        /* tamanhos[nt] */
    const int32_t* rule_len;       /* quantity of each rule; the rules of nt 0 first, then nt 1, ... */
    const int32_t* rule_symbols;   /* rule r: Symbols at [r * max_rule, r * max_rule + rule_len[r]) */
    int32_t max_rule;              /* MAX_QUANT_SIMBOLOS_POR_REGRA */
    int32_t t_const;               /* T_CONST's Symbol value, -1 without one (NOT_USING_T_CONST) */
    int32_t t_attribute;           /* T_ATTRIBUTE's Symbol value (T_ATTRk = 20000 + k decode to it) */
} ppi_ge_grammar;

/* ppi_init's options (ppi.cc:200-300). */
typedef struct ppi_ge_params {
    int32_t population_size;       /* -ps */
    int32_t number_of_bits;        /* -nb */
    int32_t bits_per_gene;         /* -bg (raised to what the grammar needs, ppi.cc:258-264) */
    int32_t bits_per_constant;     /* -bc */
    float min_constant, max_constant; /* -min, -max (swapped when min > max) */
    float crossover_rate;          /* -cp */
    float mutation_rate;           /* -mr */
    int32_t tournament_size;       /* -ts */
    int32_t elitism;               /* -e */
    int32_t max_phenotype;         /* -mps: 0 = MAX_QUANT_SIMBOLOS_POR_REGRA * nb / bg */
    int32_t threads;               /* -t: the reference's OpenMP threads = random streams (1 = sequential) */
    int64_t seed;                  /* -s; seed ^ thread must be non-zero (0 seeds from the clock) */
    float alpha;                   /* ALPHA (ppi.cc:63): 1e-6 */
} ppi_ge_params;

typedef struct ppi_ge_stats {
    int32_t generation;            /* generations evolved since ppi_ge_initialize */
    int32_t bits_per_gene;         /* effective */
    int32_t max_size_phenotype;    /* phenotype stride */
    int32_t max_slots;             /* register slots of the last compiled population */
    float best_fitness;            /* best individual so far */
    double breed_ms, decode_ms, compile_ms, evaluate_ms, rank_ms; /* last generation, device time */
    double generation_ms;          /* last generation, host wall time */
    double node_evals;             /* last generation: sum(size) * nlin */
} ppi_ge_stats;

/* Evaluates through eng's configuration and dataset; destroy it before eng. */
ppi_status ppi_ge_create(ppi_ge** out, ppi_engine* eng);
void ppi_ge_destroy(ppi_ge* ge);
ppi_status ppi_ge_configure(ppi_ge* ge, const ppi_ge_params* params, const ppi_ge_grammar* grammar);
/* ppi_generate_population (ppi.cc:656-672): random genomes, decoded and evaluated. */
ppi_status ppi_ge_initialize(ppi_ge* ge);
/* `generations` iterations of ppi_evolve's loop (ppi.cc:878-945). */
ppi_status ppi_ge_evolve(ppi_ge* ge, int32_t generations);
/* The current population as the reference holds it: genomes as GENOME_TYPE bytes (0/1, P x nb) and
   fitness (P); either may be NULL. */
ppi_status ppi_ge_get_population(const ppi_ge* ge, uint8_t* genomes, float* fitness);
/* Immigrants for the next generation (ppi_receive_individual, ppi.cc:528-560): the next
   ppi_ge_evolve generation starts its offspring population with these n genomes (bool bytes,
   n x nb), then the elite (with elitism), then the bred individuals, as ppi_evolve does with
   nImmigrants (ppi.cc:882-896). */
ppi_status ppi_ge_set_immigrants(ppi_ge* ge, const uint8_t* genomes, int32_t n);
ppi_status ppi_ge_best(const ppi_ge* ge, float* fitness, uint8_t* genome);
/* The current population decoded (ppi.cc:578-591): phenotype and ephemeral P x stride (ephemeral
   is 0 off T_CONST / T_ATTRIBUTE positions), size P; *stride = max_size_phenotype. */
ppi_status ppi_ge_decoded(const ppi_ge* ge, int32_t* phenotype, float* ephemeral, int32_t* size, int32_t* stride);
/* The device-compiled tape of the current population, laid out as ppi_engine_resident_tape. */
ppi_status ppi_ge_tape(const ppi_ge* ge, uint32_t* meta, uint32_t* code, int64_t* n_code);
ppi_status ppi_ge_get_stats(const ppi_ge* ge, ppi_ge_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* PPI_ENGINE_H */
