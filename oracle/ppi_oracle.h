/*
 * ppi_oracle.h — CPU oracle for the PPI evaluation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * algorithm (src/interpreter/sequential.cc, accelerator.cl, accelerator.cc,
 * interpreter_core, functions.h, util/Util.h) used to check the HIP engine.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The engine never links or calls it.
 *
 * Programs use the same table-driven encoding as the engine (ppi_symbol in
 * include/ppi_engine.h): a grammar Symbol value is looked up in the table to
 * find its opcode.
 */
#ifndef PPI_ORACLE_H
#define PPI_ORACLE_H

#include <stdint.h>
#include "../include/ppi_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_cfg {
    int32_t strategy;      /* ppi_strategy */
    int32_t error_fn;      /* ppi_error_fn */
    int32_t reducemax;
    int32_t protected_ops;
    int32_t cost_n;
    const float* cost;
    int32_t dp_local;      /* evaluate_dp work-group size (accelerator.cc:304-313) */
    int32_t pdp_local;     /* evaluate_pdp work-group size (accelerator.cc:316-329) */
    int32_t best_local;    /* best_individual work-group size (accelerator.cc:344-355) */
    /* PPI_EF_EXPR: the user's ERROR(X,Y) as a compiled C function of float X and Y (the
       reference pastes the expression into sequential.cc, definitions.h:18) */
    float (*ef_fn)(float X, float Y);
} oracle_cfg;

/* Error vector + ranking, exactly as the reference path named by cfg->strategy.
   data is row-major nlin x ncol with y in column ncol-1.  Returns 0, or -1 on bad input. */
int oracle_interpret(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                     const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                     int32_t stride, int32_t n_programs,
                     const float* data, int32_t nlin, int32_t ncol,
                     float alpha, float* vector, int32_t* index, int32_t* best_size,
                     int32_t ppp_mode, int32_t prediction_mode);

/* Per-row float errors of the sequential path, summed exactly: sums[i] = sum over rows of
   (double)ERROR(X,Y) with every row evaluated in float exactly as seq_interpret does
   (max instead of sum under REDUCEMAX).  bad[i] = 1 if some row's error is non-finite.
   This is the value the float sums of every path approximate; it pins reduction-order
   effects independently of the order either path uses. */
int oracle_exact_sums(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                      const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                      int32_t stride, int32_t n_programs, const float* data, int32_t nlin, int32_t ncol,
                      double* sums, int32_t* bad);

/* Value of one program on one row (float, reference semantics). */
float oracle_eval_row(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                      const int32_t* phenotype, const float* ephemeral, int32_t size,
                      const float* row, int32_t ncol);

/* util::PickNBest (Util.h:158-173): best[k] = k-th smallest error, ties -> largest index. */
void oracle_pick_n_best(int32_t n, int32_t* best, int32_t num, const float* errors, const int32_t* indices);

/* FP64 evaluation: fitness64[i] = sum_rows EF64 / nlin (or max) with every operator in double,
   non-finite handling as the sequential path.  alpha*size is NOT added. */
int oracle_fitness_f64(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                       const int32_t* phenotype, const double* ephemeral64, const int32_t* size,
                       int32_t stride, int32_t n_programs, const float* data, int32_t nlin, int32_t ncol,
                       double* fitness64);

/* FP64 forward-mode gradient of fitness64 (sum EF / nlin) with respect to every T_CONST of each
   program, in increasing phenotype position.  grad is n_programs x max_consts; n_consts[i] is
   the number of constants found (extra ones beyond max_consts are ignored). */
int oracle_grad_f64(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                    const int32_t* phenotype, const double* ephemeral64, const int32_t* size,
                    int32_t stride, int32_t n_programs, const float* data, int32_t nlin, int32_t ncol,
                    int32_t max_consts, double* fitness64, double* grad, int32_t* n_consts);
/* Same gradient, plus gmag[i][k] = sum over rows of |EF'(pred) * d pred / d const_k| / nlin: the
   magnitude of the terms the gradient sums, which bounds how far rounding of the terms can move it.
   f32_values != 0: operator values are the sequential path's float results (the tangents stay
   FP64 derivatives at those float operands), so the gradient is the exact derivative of the
   float computation; 0: everything in double. */
int oracle_grad_mag_f64(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                        const int32_t* phenotype, const double* ephemeral64, const int32_t* size,
                        int32_t stride, int32_t n_programs, const float* data, int32_t nlin, int32_t ncol,
                        int32_t max_consts, double* grad, double* gmag, int32_t f32_values);

/* The sequential path timed on `threads` host threads (threads <= 1: the reference's own
   single-threaded loop).  Returns elapsed seconds, or a negative value on error. */
double oracle_time_seq(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                       const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                       int32_t stride, int32_t n_programs, const float* data, int32_t nlin, int32_t ncol,
                       float alpha, float* vector, int32_t threads);

/* ---- grammatical evolution (ppi_ge_oracle.c) ---- */

/* XorShift128Plus of src/util/Random.h:130-162. */
typedef struct oracle_rng { uint64_t s[2]; } oracle_rng;
void oracle_rng_seed(oracle_rng* r, uint64_t seed);
uint64_t oracle_rng_int(oracle_rng* r);
double oracle_rng_real(oracle_rng* r);
int oracle_rng_probability(oracle_rng* r, double p);
/* n draws of Int() and of Real() from two generators seeded `seed`. */
int oracle_rng_draws(uint64_t seed, int32_t n, uint64_t* ints, double* reals);

/* Effective bits per gene (raised to the grammar's need) and phenotype stride (ppi.cc:258-264,300-307). */
int oracle_ge_effective(const ppi_ge_grammar* g, const ppi_ge_params* p, int32_t* bits_per_gene,
                        int32_t* max_size_phenotype);
/* decode (ppi.cc:106-182) of n genomes (bytes 0/1, n x number_of_bits): phenotype / ephemeral
   n x max_size_phenotype (0 where the reference leaves the array untouched), size n. */
int oracle_ge_decode(const ppi_ge_grammar* g, const ppi_ge_params* p, const uint8_t* genomes, int32_t n,
                     int32_t* phenotype, float* ephemeral, int32_t* size);
/* One breeding step (ppi.cc:896-936) from parents / their fitness into children[n_immigrants..P). */
int oracle_ge_breed(const ppi_ge_grammar* g, const ppi_ge_params* p, const uint8_t* parents, const float* fitness,
                    uint8_t* children, int32_t n_immigrants);
/* ppi_evolve (ppi.cc:823-945) without islands: the initial population and `generations` generations,
   each evaluated by oracle_interpret with cfg.  Returns the final population (genomes P x nb,
   fitness P), the best fitness after each generation (generations + 1 values) and the best genome. */
int oracle_ge_run(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable, const ppi_ge_grammar* g,
                  const ppi_ge_params* p, const float* data, int32_t nlin, int32_t ncol, int32_t generations,
                  uint8_t* genomes, float* fitness, float* best_history, uint8_t* best_genome);

#ifdef __cplusplus
}
#endif

#endif
