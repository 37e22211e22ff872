/*
 * ppi_interpreter.h — the drop-in interface: PPI's interpreter entry points with
 * the reference's exact C++ signatures, implemented on the MI355X engine
 * (ppi-auto_amd/shim/interpreter.cc, built into libppi_interpreter.so).
 *
 * Replaces src/interpreter/accelerator.h:29-48 and sequential.h:28-49.  Like
 * the reference, a translation unit that includes this header must see the
 * grammar's generated `symbol` header (enum Symbol) on its include path; the
 * PROFILING define selects the variant with the extra `sum_size_gen` argument
 * and the *_print_time functions, exactly as in the reference.
 */
#ifndef PPI_INTERPRETER_H
#define PPI_INTERPRETER_H

#include <symbol>

#include "ppi_individual.h"

/* accelerator.h:29 — options are read from argv (-strategy PP|DP|PDP, -cl-mls N,
   -cl-d N, -cl-p N, -type CPU|GPU, -transpose, -v); unknown options are ignored.
   Returns 0 on success, 1 on error (message on stderr). */
int acc_interpret_init(int argc, char** argv, const unsigned size, const unsigned max_arity,
                       const unsigned population_size, float** input, int nlin, int ncol, int ppp_mode,
                       int prediction_mode);

/* accelerator.h:36-40 — errors are thrown as std::runtime_error (the reference
   propagates cl::Error, a std::exception). */
void acc_interpret(Symbol* phenotype, float* ephemeral, int* size,
#ifdef PROFILING
                   unsigned long sum_size_gen,
#endif
                   float* vector, int nInd, void (*send)(Population*), int (*receive)(GENOME_TYPE**),
                   Population* migrants, int* nImmigrants, int* index, int* best_size, int ppp_mode,
                   int prediction_mode, float alpha);

#ifdef PROFILING
/* accelerator.h:48 */
void acc_print_time(bool total, unsigned long long sum_size);
#endif

/* sequential.h:28 */
void seq_interpret_init(const unsigned size, float** input, int nlin, int ncol);

/* sequential.h:35-39 */
void seq_interpret(Symbol* phenotype, float* ephemeral, int* size,
#ifdef PROFILING
                   unsigned long sum_size_gen,
#endif
                   float* vector, int nInd, int* index, int* best_size, int ppp_mode, int prediction_mode,
                   float alpha);

/* sequential.h:46 */
void seq_interpret_destroy();

#ifdef PROFILING
/* sequential.h:49 */
void seq_print_time(bool total, unsigned long long sum_size);
#endif

#endif
