/*
 * workload.h — seeded synthetic populations and datasets for benchmarks and tests.
 *
 * Programs come out in PPI's decoded phenotype layout (ppi.cc:312-314): prefix
 * order, program i in [i*stride, i*stride + size[i]), T_ATTRIBUTE with the column
 * index in ephemeral[] and T_CONST with its value in ephemeral[] (ppi.cc:155-171).
 * Datasets come out row-major, nlin x (nfeat + 1), target in the last column, as
 * PPI's CSV reader hands them to the interpreter (main.cc:32-122).
 */
#ifndef PPI_WORKLOAD_H
#define PPI_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ppi_wl_grammar {
    const int32_t* binary_ops; int32_t n_binary;  /* Symbol values of binary operators */
    const int32_t* unary_ops; int32_t n_unary;    /* Symbol values of unary operators */
    int32_t t_attribute;                          /* Symbol value of T_ATTRIBUTE */
    int32_t t_const;                              /* Symbol value of T_CONST (<0: none) */
    int32_t n_attributes;                         /* attributes attr0 .. attr{n-1} */
    float const_lo, const_hi;                     /* T_CONST ~ U[lo, hi] */
    float p_const;                                /* probability a leaf is T_CONST */
    float p_leaf;                                 /* grow: probability an inner position stops */
    int32_t max_consts;                           /* at most this many T_CONST per program (<=0: no cap) */
} ppi_wl_grammar;

/* "Grow" trees of depth <= max_depth; programs longer than max_symbols are redrawn.
   Returns the total number of symbols written, or -1 on bad arguments. */
int64_t ppi_wl_grow(const ppi_wl_grammar* g, uint64_t seed, int32_t n, int32_t stride,
                    int32_t max_depth, int32_t max_symbols,
                    int32_t* phenotype, float* ephemeral, int32_t* size);

/* Trees with an exact length drawn uniformly from [len_lo, len_hi]. */
int64_t ppi_wl_exact(const ppi_wl_grammar* g, uint64_t seed, int32_t n, int32_t stride,
                     int32_t len_lo, int32_t len_hi,
                     int32_t* phenotype, float* ephemeral, int32_t* size);

typedef enum ppi_wl_dataset_kind {
    PPI_WL_FORMULA_SMALL = 1,  /* x ~ U[-1,1]^4, y = x0^2 - 0.5*x1*x2 + sin(x3) */
    PPI_WL_POLY_NORMAL = 2,    /* x ~ N(0,1), y = random cubic polynomial + N(0, 0.01^2) */
    PPI_WL_CLUSTERS = 3,       /* 4 Gaussian clusters (sigma 0.5, centres U[-3,3]), y = cluster id */
    PPI_WL_POLY_UNIFORM = 4,   /* x ~ U[-1,1], y = random cubic polynomial */
    PPI_WL_FORMULA_ADJOINT = 5 /* x ~ N(0,1)^8, y = exp(0.5*x0) + sin(2*x1) + x2*x3*x4 */
} ppi_wl_dataset_kind;

int32_t ppi_wl_dataset(int32_t kind, uint64_t seed, int32_t nlin, int32_t nfeat, float* out);

#ifdef __cplusplus
}
#endif

#endif
