/*
 * ppi_oracle.c — CPU oracle (test infrastructure only; see ppi_oracle.h).
 *
 * Plain-C restatement of the reference's evaluation path:
 *   - opcode semantics: src/interpreter/interpreter_core (operators :1-151,
 *     named constants :152-190, integer pushes :191-793), operand order
 *     stack[top] = left operand (:48-51);
 *   - PROTECTED wrappers: src/interpreter/functions.h:59-193.  The sequential
 *     path calls the double libm routines on float arguments and rounds the
 *     result to float (C semantics of sin(float) etc.), which is what this file
 *     does by calling the double routines from C;
 *   - seq_interpret: src/interpreter/sequential.cc:117-209;
 *   - evaluate_pp / evaluate_dp / evaluate_pdp / best_individual:
 *     src/interpreter/accelerator.cl:7-313 with the host epilogues of
 *     src/interpreter/accelerator.cc:756-889;
 *   - util::PickNBest: src/util/Util.h:158-173.
 * Undefined behaviour of the reference is given one defined meaning, shared with
 * the engine: a stack underflow, an attribute column outside [0, ncol) and a
 * COST index outside the matrix all yield NaN.
 */
#define _GNU_SOURCE 1 /* exp10 */
#include "ppi_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Float constants of functions.h:19-57 / read_grammar.py symbol header. */
#define O_PI_2_F 1.57079637050629f
#define O_PI_F 3.14159274101257f
#define O_E_F 2.71828174591064f

/* ------------------------------------------------------------------------- */
/* symbol lookup                                                              */
/* ------------------------------------------------------------------------- */

typedef struct lut {
    int32_t ntable;
    const ppi_symbol* table;
    int32_t direct[4096]; /* symbol - 10000 in [0, 4096): index into table + 1 */
} lut;

static void lut_init(lut* L, const ppi_symbol* table, int32_t ntable) {
    L->ntable = ntable;
    L->table = table;
    memset(L->direct, 0, sizeof(L->direct));
    for (int32_t k = 0; k < ntable; ++k) {
        int32_t s = table[k].symbol - 10000;
        if (s >= 0 && s < 4096 && L->direct[s] == 0) L->direct[s] = k + 1;
    }
}

static const ppi_symbol* lut_find(const lut* L, int32_t symbol) {
    int32_t s = symbol - 10000;
    if (s >= 0 && s < 4096) return L->direct[s] ? &L->table[L->direct[s] - 1] : NULL;
    for (int32_t k = 0; k < L->ntable; ++k)
        if (L->table[k].symbol == symbol) return &L->table[k];
    return NULL;
}

static int op_arity(int op) {
    if (op == PPI_OP_IF_THEN_ELSE) return 3;
    if (op >= PPI_OP_AND && op <= PPI_OP_DIV) return 2;
    if (op >= PPI_OP_NOT && op <= PPI_OP_TAN) return 1;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* float semantics (sequential path)                                          */
/* ------------------------------------------------------------------------- */

static int b_(float x) { return x != 0.0f; } /* (bool)x */

/* functions.h:59-193 with NATIVE off; the CPU NATIVE shims (sequential.cc:32-74)
   compute the same values. */
static float f_SQRT(float x, int prot) { return (prot && x < 0.0f) ? 1.0f : (float)sqrt(x); }
static float f_LOG(float x, int prot) { return (prot && x < 1.0f) ? 1.0f : (float)log(x); }
static float f_LOG10(float x, int prot) { return (prot && x < 1.0f) ? 1.0f : (float)log10(x); }
static float f_LOG2(float x, int prot) { return (prot && x < 1.0f) ? 1.0f : (float)log2(x); }
static float f_DIV(float x, float y, int prot) { return (prot && y == 0.0f) ? 1.0f : x / y; }
static float f_EXP(float x) { return (float)exp(x); }
static float f_TAN(float x, int prot) {
    if (prot && (fabs(x) == O_PI_2_F || fabs(x) == 3.0f * O_PI_2_F)) return 1.0f;
    return (float)tan(x);
}

static float apply1(int op, float a, int prot) {
    switch (op) {
    case PPI_OP_NOT: return (float)!b_(a);
    case PPI_OP_ABS: return (float)fabs(a);
    case PPI_OP_POW2: return a * a;
    case PPI_OP_POW3: return a * a * a;
    case PPI_OP_POW4: return a * a * a * a;
    case PPI_OP_POW5: return a * a * a * a * a;
    case PPI_OP_NEG: return -a;
    case PPI_OP_ROUND: return (float)round(a);
    case PPI_OP_CEIL: return (float)ceil(a);
    case PPI_OP_FLOOR: return (float)floor(a);
    case PPI_OP_EXP: return f_EXP(a);
    case PPI_OP_EXP10: return (float)exp10(a);
    case PPI_OP_EXP2: return (float)exp2(a);
    case PPI_OP_SIN: return (float)sin(a);
    case PPI_OP_COS: return (float)cos(a);
    case PPI_OP_STEP: return (float)(a >= 0.0f);
    case PPI_OP_SIGN: return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f);
    case PPI_OP_LOGISTIC: return 1.0f / (1.0f + f_EXP(-a));
    case PPI_OP_GAMMA: {
        /* interpreter_core:127-129; sinh and pow are the double routines, so the
           products that involve them are formed in double before rounding. */
        float base = (a / O_E_F) * f_SQRT((float)((double)a * sinh(1.0f / a)), prot);
        return (float)(pow(base, a) * (double)f_SQRT(2.0f * O_PI_F / a, prot));
    }
    case PPI_OP_GAUSS: return f_EXP(-a * a);
    case PPI_OP_SQRT: return f_SQRT(a, prot);
    case PPI_OP_LOG: return f_LOG(a, prot);
    case PPI_OP_LOG10: return f_LOG10(a, prot);
    case PPI_OP_LOG2: return f_LOG2(a, prot);
    case PPI_OP_TAN: return f_TAN(a, prot);
    default: return NAN;
    }
}

/* a = stack[top] (left operand), b = stack[top-1]. */
static float apply2(int op, float a, float b, int prot) {
    switch (op) {
    case PPI_OP_AND: return (b_(a) && b_(b)) ? 1.0f : 0.0f;
    case PPI_OP_OR: return (b_(a) || b_(b)) ? 1.0f : 0.0f;
    case PPI_OP_XOR: return (!b_(a) != !b_(b)) ? 1.0f : 0.0f;
    case PPI_OP_GREATER: return a > b ? 1.0f : 0.0f;
    case PPI_OP_GREATEREQUAL: return a >= b ? 1.0f : 0.0f;
    case PPI_OP_LESS: return a < b ? 1.0f : 0.0f;
    case PPI_OP_LESSEQUAL: return a <= b ? 1.0f : 0.0f;
    case PPI_OP_EQUAL: return a == b ? 1.0f : 0.0f;
    case PPI_OP_NOTEQUAL: return a != b ? 1.0f : 0.0f;
    case PPI_OP_ADD: return a + b;
    case PPI_OP_SUB: return a - b;
    case PPI_OP_MULT: return a * b;
    case PPI_OP_MEAN: return (a + b) / 2.0f;
    case PPI_OP_MAX: return (float)fmax(a, b);
    case PPI_OP_MIN: return (float)fmin(a, b);
    case PPI_OP_MOD: return (float)fmod(a, b);
    case PPI_OP_POW: return (float)pow(a, b);
    case PPI_OP_DIV: return f_DIV(a, b, prot);
    default: return NAN;
    }
}

/* One program on one row; the stack scan of sequential.cc:143-164. */
static float eval_f32(const lut* L, int prot, const int32_t* ph, const float* eph, int32_t size,
                      const float* row, int32_t ncol, float* stack) {
    int top = -1;
    for (int32_t i = size - 1; i >= 0; --i) {
        const ppi_symbol* s = lut_find(L, ph[i]);
        int op = s ? s->op : PPI_OP_UNKNOWN;
        int ar = op_arity(op);
        if (ar > 0 && top < ar - 1) return NAN; /* underflow: undefined in the reference */
        switch (ar) {
        case 3:
            stack[top - 2] = b_(stack[top]) ? stack[top - 1] : stack[top - 2];
            top -= 2;
            break;
        case 2:
            stack[top - 1] = apply2(op, stack[top], stack[top - 1], prot);
            --top;
            break;
        case 1:
            stack[top] = apply1(op, stack[top], prot);
            break;
        default:
            if (op == PPI_OP_CONST) {
                stack[++top] = eph[i];
            } else if (op == PPI_OP_ATTRIBUTE) {
                float c = eph[i];
                stack[++top] = (c > -1.0f && c < (float)ncol) ? row[(int)c] : NAN;
            } else if (op == PPI_OP_VALUE) {
                stack[++top] = s->value;
            } else {
                stack[++top] = NAN;
            }
        }
    }
    return top >= 0 ? stack[top] : NAN;
}

/* ERROR(X,Y) variants (definitions.h:18). */
static float ef_f32(const oracle_cfg* c, float X, float Y) {
    switch (c->error_fn) {
    case PPI_EF_ABS: return (float)fabs(X - Y);
    case PPI_EF_SQUARED: return (X - Y) * (X - Y);
    case PPI_EF_NEQ: return (float)(X != Y);
    case PPI_EF_COST:
    case PPI_EF_COST_CLAMP: {
        int n = c->cost_n;
        float xx = X;
        if (c->error_fn == PPI_EF_COST_CLAMP) xx = fminf(fmaxf(X, 0.0f), (float)(n - 1));
        if (!(xx > -1.0f && xx < (float)n) || !(Y > -1.0f && Y < (float)n)) return NAN;
        return c->cost[(int)xx * n + (int)Y];
    }
    case PPI_EF_EXPR: return c->ef_fn ? c->ef_fn(X, Y) : NAN;
    default: return NAN;
    }
}

static int fin(float x) { return !(isinf(x) || isnan(x)); }

/* ------------------------------------------------------------------------- */
/* ranking                                                                    */
/* ------------------------------------------------------------------------- */

typedef struct ei { float e; int32_t i; } ei;

static int cmp_ei(const void* pa, const void* pb) {
    const ei* a = (const ei*)pa;
    const ei* b = (const ei*)pb;
    if (a->e < b->e) return -1;
    if (b->e < a->e) return 1;
    return (a->i < b->i) ? 1 : (a->i > b->i ? -1 : 0); /* ties: larger index first */
}

void oracle_pick_n_best(int32_t n, int32_t* best, int32_t num, const float* errors, const int32_t* indices) {
    /* priority_queue<pair<float,int>> of (-error, i): pops smallest error first,
       and on equal errors the largest i first. */
    ei* v = (ei*)malloc(sizeof(ei) * (size_t)(num > 0 ? num : 1));
    for (int32_t i = 0; i < num; ++i) { v[i].e = errors[i]; v[i].i = i; }
    qsort(v, (size_t)num, sizeof(ei), cmp_ei);
    for (int32_t k = 0; k < n && k < num; ++k) best[k] = indices ? indices[v[k].i] : v[k].i;
    free(v);
}

/* pown(2, ceil(log2(L))) computed exactly. */
static int npo2(int L) { int p = 1; while (p < L) p <<= 1; return p; }

/* best_individual (accelerator.cl:270-313) + host PickNBest (accelerator.cc:866-883). */
static void best_acc(const float* vec, int32_t P, int32_t local2, int32_t* index, int32_t* best_size) {
    int32_t ng = (P + local2 - 1) / local2;
    float* PB = (float*)malloc(sizeof(float) * (size_t)ng);
    int32_t* PI = (int32_t*)malloc(sizeof(int32_t) * (size_t)ng);
    float* lb = (float*)malloc(sizeof(float) * (size_t)local2);
    int32_t* li = (int32_t*)malloc(sizeof(int32_t) * (size_t)local2);
    for (int32_t g = 0; g < ng; ++g) {
        for (int32_t l = 0; l < local2; ++l) {
            int32_t gl = g * local2 + l;
            if (gl < P) { lb[l] = vec[gl]; li[l] = gl; }
            else { lb[l] = FLT_MAX; li[l] = -1; }
        }
        for (int s = npo2(local2) / 2; s > 0; s /= 2)
            for (int l = 0; l < s; ++l)
                if (l + s < local2 && lb[l] > lb[l + s]) { lb[l] = lb[l + s]; li[l] = li[l + s]; }
        PB[g] = lb[0];
        PI[g] = li[0];
    }
    if (*best_size > ng) *best_size = ng;
    oracle_pick_n_best(*best_size, index, ng, PB, PI);
    free(PB); free(PI); free(lb); free(li);
}

/* ------------------------------------------------------------------------- */
/* the four reference paths                                                   */
/* ------------------------------------------------------------------------- */

int oracle_interpret(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                     const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                     int32_t stride, int32_t nInd, const float* data, int32_t nlin, int32_t ncol,
                     float alpha, float* vector, int32_t* index, int32_t* best_size,
                     int32_t ppp_mode, int32_t prediction_mode) {
    if (!cfg || nlin <= 0 || ncol <= 0 || nInd < 0 || stride <= 0) return -1;
    lut L;
    lut_init(&L, table, ntable);
    const int prot = cfg->protected_ops;
    float* stack = (float*)malloc(sizeof(float) * (size_t)(stride + 1));
    const float fn = (float)nlin;

    if (ppp_mode && prediction_mode) {
        for (int32_t ind = 0; ind < nInd; ++ind)
            for (int32_t r = 0; r < nlin; ++r)
                vector[r] = eval_f32(&L, prot, phenotype + (size_t)ind * stride, ephemeral + (size_t)ind * stride,
                                     size[ind], data + (size_t)r * ncol, ncol, stack);
        free(stack);
        return 0;
    }

    float* raw = (float*)malloc(sizeof(float) * (size_t)(nInd > 0 ? nInd : 1));

    for (int32_t ind = 0; ind < nInd; ++ind) {
        const int32_t* ph = phenotype + (size_t)ind * stride;
        const float* eph = ephemeral + (size_t)ind * stride;
        const int32_t sz = size[ind];
        if (cfg->strategy == PPI_STRATEGY_SEQ || cfg->strategy == PPI_STRATEGY_PP) {
            /* sequential.cc:133-190 / accelerator.cl:22-84: rows in order, break at the
               first non-finite error. */
            if (sz == 0) {
                raw[ind] = FLT_MAX;
                vector[ind] = FLT_MAX;
                continue;
            }
            float sum = 0.0f;
            for (int32_t r = 0; r < nlin; ++r) {
                const float* row = data + (size_t)r * ncol;
                float err = ef_f32(cfg, eval_f32(&L, prot, ph, eph, sz, row, ncol, stack), row[ncol - 1]);
                if (!fin(err)) { sum = FLT_MAX; break; }
                if (cfg->reducemax) sum = (err * fn > sum) ? err * fn : sum;
                else sum += err;
            }
            if (cfg->strategy == PPI_STRATEGY_SEQ) {
                vector[ind] = fin(sum) ? sum / fn + alpha * (float)sz : FLT_MAX;
            } else {
                raw[ind] = fin(sum) ? sum / fn : FLT_MAX;
                vector[ind] = raw[ind] + alpha * (float)sz;
            }
        } else if (cfg->strategy == PPI_STRATEGY_FP) {
            /* evaluate_dp (accelerator.cl:88-171) + host reduction (accelerator.cc:773-805). */
            const int32_t Ld = cfg->dp_local;
            const int32_t ng = (nlin + Ld - 1) / Ld;
            float* PE = (float*)malloc(sizeof(float) * (size_t)Ld);
            float sum = 0.0f;
            for (int32_t g = 0; g < ng; ++g) {
                float part;
                if (sz == 0) {
                    part = FLT_MAX;
                } else {
                    for (int32_t l = 0; l < Ld; ++l) {
                        int32_t r = g * Ld + l;
                        PE[l] = 0.0f;
                        if (r < nlin) {
                            const float* row = data + (size_t)r * ncol;
                            PE[l] = ef_f32(cfg, eval_f32(&L, prot, ph, eph, sz, row, ncol, stack), row[ncol - 1]);
                        }
                    }
                    for (int s = npo2(Ld) / 2; s > 0; s /= 2)
                        for (int l = 0; l < s; ++l)
                            if (l + s < Ld) {
                                if (cfg->reducemax) PE[l] = (PE[l + s] > PE[l]) ? PE[l + s] : PE[l];
                                else PE[l] += PE[l + s];
                            }
                    part = PE[0];
                }
                if (!fin(part)) { sum = FLT_MAX; break; }
                if (cfg->reducemax) sum = (part * fn > sum) ? part * fn : sum;
                else sum += part;
            }
            free(PE);
            vector[ind] = fin(sum) ? sum / fn + alpha * (float)sz : FLT_MAX;
            raw[ind] = vector[ind];
        } else if (cfg->strategy == PPI_STRATEGY_PDP) {
            /* evaluate_pdp (accelerator.cl:173-268) + host epilogue (accelerator.cc:849). */
            if (sz == 0) {
                raw[ind] = FLT_MAX;
            } else {
                const int32_t Lp = cfg->pdp_local;
                float* PE = (float*)malloc(sizeof(float) * (size_t)Lp);
                const int32_t nj = (nlin + Lp - 1) / Lp;
                for (int32_t l = 0; l < Lp; ++l) {
                    PE[l] = 0.0f;
                    for (int32_t j = 0; j < nj; ++j) {
                        int32_t r = j * Lp + l;
                        if (r >= nlin) continue;
                        const float* row = data + (size_t)r * ncol;
                        float err = ef_f32(cfg, eval_f32(&L, prot, ph, eph, sz, row, ncol, stack), row[ncol - 1]);
                        if (!fin(err)) { PE[l] = FLT_MAX; break; }
                        if (cfg->reducemax) PE[l] = (err * fn > PE[l]) ? err * fn : PE[l];
                        else PE[l] += err;
                    }
                }
                for (int s = npo2(Lp) / 2; s > 0; s >>= 1)
                    for (int l = 0; l < s; ++l)
                        if (l + s < Lp) {
                            if (cfg->reducemax) PE[l] = (PE[l + s] > PE[l]) ? PE[l + s] : PE[l];
                            else PE[l] += PE[l + s];
                        }
                raw[ind] = fin(PE[0]) ? PE[0] / fn : FLT_MAX;
                free(PE);
            }
            vector[ind] = raw[ind] + alpha * (float)sz;
        } else {
            free(raw); free(stack);
            return -1;
        }
    }

    if (!ppp_mode && index && best_size) {
        if (cfg->strategy == PPI_STRATEGY_SEQ) {
            /* sequential.cc:197-208 */
            if (*best_size > nInd) *best_size = nInd;
            oracle_pick_n_best(*best_size, index, nInd, vector, NULL);
        } else {
            /* PP/PDP rank the kernel output (no alpha); DP ranks the alpha-inclusive vector. */
            best_acc(raw, nInd, cfg->best_local, index, best_size);
        }
    }
    free(raw);
    free(stack);
    return 0;
}

int oracle_exact_sums(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                      const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                      int32_t stride, int32_t nInd, const float* data, int32_t nlin, int32_t ncol,
                      double* sums, int32_t* bad) {
    if (!cfg || nlin <= 0 || ncol <= 0 || stride <= 0) return -1;
    lut L;
    lut_init(&L, table, ntable);
#pragma omp parallel for schedule(dynamic, 8)
    for (int32_t ind = 0; ind < nInd; ++ind) {
        float* stack = (float*)malloc(sizeof(float) * (size_t)(stride + 1));
        const int32_t* ph = phenotype + (size_t)ind * stride;
        const float* eph = ephemeral + (size_t)ind * stride;
        double acc = 0.0;
        int b = size[ind] == 0;
        for (int32_t r = 0; r < nlin && !b; ++r) {
            const float* row = data + (size_t)r * ncol;
            float err = ef_f32(cfg, eval_f32(&L, cfg->protected_ops, ph, eph, size[ind], row, ncol, stack),
                               row[ncol - 1]);
            if (!fin(err)) { b = 1; break; }
            if (cfg->reducemax) { double v = (double)(err * (float)nlin); acc = v > acc ? v : acc; }
            else acc += (double)err;
        }
        sums[ind] = acc;
        bad[ind] = b;
        free(stack);
    }
    return 0;
}

float oracle_eval_row(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                      const int32_t* phenotype, const float* ephemeral, int32_t size,
                      const float* row, int32_t ncol) {
    lut L;
    lut_init(&L, table, ntable);
    float* stack = (float*)malloc(sizeof(float) * (size_t)(size + 1));
    float v = eval_f32(&L, cfg->protected_ops, phenotype, ephemeral, size, row, ncol, stack);
    free(stack);
    return v;
}

/* ------------------------------------------------------------------------- */
/* FP64 evaluation and forward-mode gradient                                  */
/* ------------------------------------------------------------------------- */

#define MAXK 16

typedef struct dual { double v; double d[MAXK]; } dual;

static double d_bool(double x) { return x != 0.0; }

/* Value and derivative of one operator in double.  K = number of tangent
   components carried (0: value only). */
static dual dapply(int op, const dual* a, const dual* b, const dual* c, int K, int prot) {
    dual r;
    memset(&r, 0, sizeof(r));
    double x = a->v, y = b ? b->v : 0.0;
    int k;
#define DLOOP(expr) for (k = 0; k < K; ++k) r.d[k] = (expr)
    switch (op) {
    case PPI_OP_IF_THEN_ELSE:
        r = d_bool(x) ? *b : *c;
        return r;
    case PPI_OP_AND: r.v = (d_bool(x) && d_bool(y)) ? 1.0 : 0.0; return r;
    case PPI_OP_OR: r.v = (d_bool(x) || d_bool(y)) ? 1.0 : 0.0; return r;
    case PPI_OP_XOR: r.v = (!d_bool(x) != !d_bool(y)) ? 1.0 : 0.0; return r;
    case PPI_OP_GREATER: r.v = x > y; return r;
    case PPI_OP_GREATEREQUAL: r.v = x >= y; return r;
    case PPI_OP_LESS: r.v = x < y; return r;
    case PPI_OP_LESSEQUAL: r.v = x <= y; return r;
    case PPI_OP_EQUAL: r.v = x == y; return r;
    case PPI_OP_NOTEQUAL: r.v = x != y; return r;
    case PPI_OP_ADD: r.v = x + y; DLOOP(a->d[k] + b->d[k]); return r;
    case PPI_OP_SUB: r.v = x - y; DLOOP(a->d[k] - b->d[k]); return r;
    case PPI_OP_MULT: r.v = x * y; DLOOP(a->d[k] * y + x * b->d[k]); return r;
    case PPI_OP_MEAN: r.v = (x + y) / 2.0; DLOOP((a->d[k] + b->d[k]) / 2.0); return r;
    case PPI_OP_MAX: r = (isnan(y) || x >= y) ? *a : *b; r.v = fmax(x, y); return r;
    case PPI_OP_MIN: r = (isnan(y) || x <= y) ? *a : *b; r.v = fmin(x, y); return r;
    case PPI_OP_MOD: { double q = trunc(x / y); r.v = fmod(x, y); DLOOP(a->d[k] - q * b->d[k]); return r; }
    case PPI_OP_POW: {
        r.v = pow(x, y);
        double lx = x > 0.0 ? log(x) : 0.0;
        DLOOP(y * pow(x, y - 1.0) * a->d[k] + (b->d[k] != 0.0 ? r.v * lx * b->d[k] : 0.0));
        return r;
    }
    case PPI_OP_DIV:
        if (prot && y == 0.0) { r.v = 1.0; return r; }
        r.v = x / y; DLOOP((a->d[k] * y - x * b->d[k]) / (y * y)); return r;
    case PPI_OP_NOT: r.v = !d_bool(x); return r;
    case PPI_OP_ABS: r.v = fabs(x); DLOOP(x > 0 ? a->d[k] : (x < 0 ? -a->d[k] : 0.0)); return r;
    case PPI_OP_POW2: r.v = x * x; DLOOP(2.0 * x * a->d[k]); return r;
    case PPI_OP_POW3: r.v = x * x * x; DLOOP(3.0 * x * x * a->d[k]); return r;
    case PPI_OP_POW4: r.v = x * x * x * x; DLOOP(4.0 * x * x * x * a->d[k]); return r;
    case PPI_OP_POW5: r.v = x * x * x * x * x; DLOOP(5.0 * x * x * x * x * a->d[k]); return r;
    case PPI_OP_NEG: r.v = -x; DLOOP(-a->d[k]); return r;
    case PPI_OP_ROUND: r.v = round(x); return r;
    case PPI_OP_CEIL: r.v = ceil(x); return r;
    case PPI_OP_FLOOR: r.v = floor(x); return r;
    case PPI_OP_EXP: r.v = exp(x); DLOOP(r.v * a->d[k]); return r;
    case PPI_OP_EXP10: r.v = exp10(x); DLOOP(r.v * M_LN10 * a->d[k]); return r;
    case PPI_OP_EXP2: r.v = exp2(x); DLOOP(r.v * M_LN2 * a->d[k]); return r;
    case PPI_OP_SIN: r.v = sin(x); DLOOP(cos(x) * a->d[k]); return r;
    case PPI_OP_COS: r.v = cos(x); DLOOP(-sin(x) * a->d[k]); return r;
    case PPI_OP_STEP: r.v = x >= 0.0; return r;
    case PPI_OP_SIGN: r.v = x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); return r;
    case PPI_OP_LOGISTIC: { double s = 1.0 / (1.0 + exp(-x)); r.v = s; DLOOP(s * (1.0 - s) * a->d[k]); return r; }
    case PPI_OP_GAMMA: {
        double e = (double)O_E_F, pi = (double)O_PI_F;
        double in1 = x * sinh(1.0 / x), in2 = 2.0 * pi / x;
        int p1 = prot && in1 < 0.0, p2 = prot && in2 < 0.0;
        double s1 = p1 ? 1.0 : sqrt(in1), s2 = p2 ? 1.0 : sqrt(in2);
        double B = (x / e) * s1;
        r.v = pow(B, x) * s2;
        double ds1 = p1 ? 0.0 : 0.5 / s1 * (sinh(1.0 / x) - cosh(1.0 / x) / x);
        double ds2 = p2 ? 0.0 : 0.5 / s2 * (-2.0 * pi / (x * x));
        double dB = s1 / e + (x / e) * ds1;
        double dlog = log(B) + x * dB / B + ds2 / s2;
        DLOOP(r.v * dlog * a->d[k]);
        return r;
    }
    case PPI_OP_GAUSS: r.v = exp(-x * x); DLOOP(-2.0 * x * r.v * a->d[k]); return r;
    case PPI_OP_SQRT:
        if (prot && x < 0.0) { r.v = 1.0; return r; }
        r.v = sqrt(x); DLOOP(0.5 / r.v * a->d[k]); return r;
    case PPI_OP_LOG:
        if (prot && x < 1.0) { r.v = 1.0; return r; }
        r.v = log(x); DLOOP(a->d[k] / x); return r;
    case PPI_OP_LOG10:
        if (prot && x < 1.0) { r.v = 1.0; return r; }
        r.v = log10(x); DLOOP(a->d[k] / (x * M_LN10)); return r;
    case PPI_OP_LOG2:
        if (prot && x < 1.0) { r.v = 1.0; return r; }
        r.v = log2(x); DLOOP(a->d[k] / (x * M_LN2)); return r;
    case PPI_OP_TAN:
        if (prot && (fabs(x) == (double)O_PI_2_F || fabs(x) == (double)(3.0f * O_PI_2_F))) { r.v = 1.0; return r; }
        r.v = tan(x); DLOOP((1.0 + r.v * r.v) * a->d[k]); return r;
    default:
        r.v = NAN;
        return r;
    }
#undef DLOOP
}

/* Prefix program in double with K tangent components; constant j (in increasing
   position order) seeds component j.  f32v: every operator's value is replaced by
   the sequential path's float result (apply1/apply2) while its tangent stays the
   exact derivative at the float operands, i.e. the derivative of the computation
   the float evaluators perform, free of float-vs-double forward differences. */
static dual eval_dual(const lut* L, int prot, const int32_t* ph, const double* eph, int32_t size,
                      const float* row, int32_t ncol, int K, dual* stack, const int32_t* const_slot, int f32v) {
    int top = -1;
    for (int32_t i = size - 1; i >= 0; --i) {
        const ppi_symbol* s = lut_find(L, ph[i]);
        int op = s ? s->op : PPI_OP_UNKNOWN;
        int ar = op_arity(op);
        if (ar > 0 && top < ar - 1) { dual n; memset(&n, 0, sizeof(n)); n.v = NAN; return n; }
        if (ar == 3) {
            stack[top - 2] = dapply(op, &stack[top], &stack[top - 1], &stack[top - 2], K, prot);
            top -= 2;
        } else if (ar == 2) {
            const double x = stack[top].v, y = stack[top - 1].v;
            stack[top - 1] = dapply(op, &stack[top], &stack[top - 1], NULL, K, prot);
            if (f32v) stack[top - 1].v = (double)apply2(op, (float)x, (float)y, prot);
            --top;
        } else if (ar == 1) {
            const double x = stack[top].v;
            stack[top] = dapply(op, &stack[top], NULL, NULL, K, prot);
            if (f32v) stack[top].v = (double)apply1(op, (float)x, prot);
        } else {
            dual* t = &stack[++top];
            memset(t, 0, sizeof(*t));
            if (op == PPI_OP_CONST) {
                t->v = eph[i];
                if (const_slot && const_slot[i] >= 0 && const_slot[i] < K) t->d[const_slot[i]] = 1.0;
            } else if (op == PPI_OP_ATTRIBUTE) {
                double c = eph[i];
                t->v = (c > -1.0 && c < (double)ncol) ? (double)row[(int)c] : NAN;
            } else if (op == PPI_OP_VALUE) {
                t->v = (double)s->value;
            } else {
                t->v = NAN;
            }
        }
    }
    if (top < 0) { dual n; memset(&n, 0, sizeof(n)); n.v = NAN; return n; }
    return stack[top];
}

static double ef_f64(const oracle_cfg* c, double X, double Y, double* dX) {
    *dX = 0.0;
    switch (c->error_fn) {
    case PPI_EF_ABS: *dX = X > Y ? 1.0 : (X < Y ? -1.0 : 0.0); return fabs(X - Y);
    case PPI_EF_SQUARED: *dX = 2.0 * (X - Y); return (X - Y) * (X - Y);
    case PPI_EF_NEQ: return X != Y;
    case PPI_EF_COST:
    case PPI_EF_COST_CLAMP: {
        int n = c->cost_n;
        double xx = X;
        if (c->error_fn == PPI_EF_COST_CLAMP) xx = fmin(fmax(X, 0.0), (double)(n - 1));
        if (!(xx > -1.0 && xx < (double)n) || !(Y > -1.0 && Y < (double)n)) return NAN;
        return c->cost[(int)xx * n + (int)Y];
    }
    default: return NAN;
    }
}

static int grad_impl(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                     const int32_t* phenotype, const double* eph64, const int32_t* size,
                     int32_t stride, int32_t nInd, const float* data, int32_t nlin, int32_t ncol,
                     int32_t max_consts, double* fitness64, double* grad, int32_t* n_consts, double* gmag,
                     int f32v) {
    if (!cfg || nlin <= 0 || ncol <= 0 || stride <= 0 || max_consts > MAXK) return -1;
    lut L;
    lut_init(&L, table, ntable);
    int fail = 0;
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t ind = 0; ind < nInd; ++ind) {
        const int32_t* ph = phenotype + (size_t)ind * stride;
        const double* eph = eph64 + (size_t)ind * stride;
        const int32_t sz = size[ind];
        dual* stack = (dual*)malloc(sizeof(dual) * (size_t)(sz + 1));
        int32_t* slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)(sz + 1));
        if (!stack || !slot) { fail = 1; free(stack); free(slot); continue; }
        int nc = 0;
        for (int32_t i = 0; i < sz; ++i) {
            const ppi_symbol* s = lut_find(&L, ph[i]);
            slot[i] = -1;
            if (s && s->op == PPI_OP_CONST) { slot[i] = nc < max_consts ? nc : -1; ++nc; }
        }
        int K = nc < max_consts ? nc : max_consts;
        if (n_consts) n_consts[ind] = nc;
        double g[MAXK] = {0}, gm[MAXK] = {0};
        double acc = 0.0;
        int bad = (sz == 0);
        double best = -1.0;
        for (int32_t r = 0; r < nlin && !bad; ++r) {
            const float* row = data + (size_t)r * ncol;
            dual v = eval_dual(&L, cfg->protected_ops, ph, eph, sz, row, ncol, K, stack, slot, f32v);
            double dX;
            double e = ef_f64(cfg, v.v, (double)row[ncol - 1], &dX);
            if (isinf(e) || isnan(e)) { bad = 1; break; }
            if (cfg->reducemax) {
                if (e > best) { best = e; for (int k = 0; k < K; ++k) g[k] = dX * v.d[k]; }
            } else {
                acc += e;
                for (int k = 0; k < K; ++k) { g[k] += dX * v.d[k]; gm[k] += fabs(dX * v.d[k]); }
            }
        }
        if (fitness64) fitness64[ind] = bad ? INFINITY : (cfg->reducemax ? best : acc / (double)nlin);
        if (grad)
            for (int k = 0; k < max_consts; ++k)
                grad[(size_t)ind * max_consts + k] =
                    bad ? NAN : (k < K ? (cfg->reducemax ? g[k] : g[k] / (double)nlin) : 0.0);
        if (gmag)
            for (int k = 0; k < max_consts; ++k)
                gmag[(size_t)ind * max_consts + k] = (bad || k >= K) ? 0.0 : gm[k] / (double)nlin;
        free(stack);
        free(slot);
    }
    return fail ? -1 : 0;
}

int oracle_fitness_f64(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                       const int32_t* phenotype, const double* ephemeral64, const int32_t* size,
                       int32_t stride, int32_t n_programs, const float* data, int32_t nlin, int32_t ncol,
                       double* fitness64) {
    return grad_impl(cfg, table, ntable, phenotype, ephemeral64, size, stride, n_programs, data, nlin, ncol,
                     0, fitness64, NULL, NULL, NULL, 0);
}

int oracle_grad_f64(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                    const int32_t* phenotype, const double* ephemeral64, const int32_t* size,
                    int32_t stride, int32_t n_programs, const float* data, int32_t nlin, int32_t ncol,
                    int32_t max_consts, double* fitness64, double* grad, int32_t* n_consts) {
    return grad_impl(cfg, table, ntable, phenotype, ephemeral64, size, stride, n_programs, data, nlin, ncol,
                     max_consts, fitness64, grad, n_consts, NULL, 0);
}

int oracle_grad_mag_f64(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                        const int32_t* phenotype, const double* ephemeral64, const int32_t* size,
                        int32_t stride, int32_t n_programs, const float* data, int32_t nlin, int32_t ncol,
                        int32_t max_consts, double* grad, double* gmag, int32_t f32_values) {
    return grad_impl(cfg, table, ntable, phenotype, ephemeral64, size, stride, n_programs, data, nlin, ncol,
                     max_consts, NULL, grad, NULL, gmag, f32_values);
}

/* ------------------------------------------------------------------------- */
/* CPU baseline timing                                                        */
/* ------------------------------------------------------------------------- */

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double oracle_time_seq(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable,
                       const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                       int32_t stride, int32_t nInd, const float* data, int32_t nlin, int32_t ncol,
                       float alpha, float* vector, int32_t threads) {
    if (threads <= 1) {
        double t0 = now_s();
        oracle_cfg c = *cfg;
        c.strategy = PPI_STRATEGY_SEQ;
        if (oracle_interpret(&c, table, ntable, phenotype, ephemeral, size, stride, nInd, data, nlin, ncol,
                             alpha, vector, NULL, NULL, 1, 0) != 0)
            return -1.0;
        return now_s() - t0;
    }
#ifdef _OPENMP
    double t0 = now_s();
    int fail = 0;
    oracle_cfg c = *cfg;
    c.strategy = PPI_STRATEGY_SEQ;
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads)
    for (int32_t ind = 0; ind < nInd; ++ind) {
        if (oracle_interpret(&c, table, ntable, phenotype + (size_t)ind * stride, ephemeral + (size_t)ind * stride,
                             size + ind, stride, 1, data, nlin, ncol, alpha, vector + ind, NULL, NULL, 1, 0) != 0)
            fail = 1;
    }
    return fail ? -1.0 : now_s() - t0;
#else
    (void)threads;
    return -2.0;
#endif
}
