/*
 * ppi_ge_oracle.c — CPU oracle of PPI's grammatical-evolution loop (test infrastructure only;
 * see ppi_oracle.h).
 *
 * Plain-C restatement of:
 *   - XorShift128Plus and BaseRNG::Real / Probability: src/util/Random.h:60-113,130-162;
 *   - one generator per OpenMP thread, seeded seed ^ thread: src/ppi.cc:388-396;
 *   - decode_rule / decode_real / decode: src/ppi.cc:106-182;
 *   - ppi_generate_population: src/ppi.cc:656-672;
 *   - ppi_tournament: src/ppi.cc:422-452;
 *   - ppi_crossover / ppi_mutation / ppi_clone: src/ppi.cc:400-418,680-814;
 *   - ppi_evolve's generation loop and ppi_evaluate's best tracking: src/ppi.cc:570-640,823-945;
 *   - the bits-per-gene and phenotype-size rules of ppi_init: src/ppi.cc:258-264,300-307.
 * OpenMP's static schedule (what `#pragma omp parallel for` uses without a schedule clause) gives
 * thread t of T the iterations [q*t + r', q*t + r' + q) with q = n / T (+1 for the first n % T
 * threads); each thread draws from its own generator in iteration order, so a run with T threads
 * is reproduced by running the T streams one after another.
 * Undefined behaviour of the reference is given one meaning, shared with the engine: a terminal
 * written at a position >= max_size_phenotype (possible only when -mps caps the size) kills the
 * individual (size 0), like the position check of decode (ppi.cc:143); a seed ^ thread of 0 (which
 * the reference replaces by the clock) is rejected; a draw that rounds to 1.0 (probability 2^-54)
 * is clamped: a random index to n - 1, the number of mutated bits to its maximum.
 */
#include "ppi_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* XorShift128Plus (Random.h:130-162)                                          */
/* ------------------------------------------------------------------------- */

void oracle_rng_seed(oracle_rng* r, uint64_t seed) {
    /* StoreSeed: both words (seed << 32) | (seed & 0xffffffff) */
    const uint64_t v = (seed << 32) | (seed & 0xffffffffull);
    r->s[0] = v;
    r->s[1] = v;
}

uint64_t oracle_rng_int(oracle_rng* r) {
    uint64_t x = r->s[0];
    const uint64_t y = r->s[1];
    r->s[0] = y;
    x ^= x << 23;
    r->s[1] = x ^ y ^ (x >> 17) ^ (y >> 26);
    return r->s[1] + y;
}

/* Real(): double(Int()) / (double(rand_max) + 1.0), rand_max = 2^64 - 1 (its double is 2^64). */
double oracle_rng_real(oracle_rng* r) {
    return (double)oracle_rng_int(r) / ((double)18446744073709551615ull + 1.0);
}

/* Probability(p): no draw when p <= 0 or p >= 1. */
int oracle_rng_probability(oracle_rng* r, double p) {
    if (p <= 0.0) return 0;
    if (p >= 1.0) return 1;
    return oracle_rng_real(r) < p;
}

int oracle_rng_draws(uint64_t seed, int32_t n, uint64_t* ints, double* reals) {
    oracle_rng a, b;
    oracle_rng_seed(&a, seed);
    oracle_rng_seed(&b, seed);
    for (int32_t i = 0; i < n; ++i) {
        if (ints) ints[i] = oracle_rng_int(&a);
        if (reals) reals[i] = oracle_rng_real(&b);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* parameters (ppi.cc:258-264, 300-307)                                        */
/* ------------------------------------------------------------------------- */

int oracle_ge_effective(const ppi_ge_grammar* g, const ppi_ge_params* p, int32_t* bits_per_gene,
                        int32_t* max_size_phenotype) {
    if (!g || !p || g->n_nonterminals <= 0 || g->max_rule <= 0 || p->number_of_bits <= 0) return -1;
    int bg = p->bits_per_gene;
    for (int i = 0; i < g->n_nonterminals; ++i) {
        const int need = (int)ceil(log10((double)g->n_rules[i]) / log10(2.0));
        if (need > bg) bg = need;
    }
    int64_t mx = (int64_t)g->max_rule * p->number_of_bits / bg;
    if (p->max_phenotype > 0 && p->max_phenotype < mx) mx = p->max_phenotype;
    if (bits_per_gene) *bits_per_gene = bg;
    if (max_size_phenotype) *max_size_phenotype = (int32_t)mx;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* decode (ppi.cc:106-182)                                                     */
/* ------------------------------------------------------------------------- */

typedef struct ge_ctx {
    const ppi_ge_grammar* g;
    int32_t* first_rule; /* per non-terminal: index of its first rule */
    int nb, bg, bc, max_size;
    float lo, hi;
} ge_ctx;

/* Returns the rule index, or -1 (not enough bits). */
static int decode_rule(const ge_ctx* c, const uint8_t* genome, int* allele, int head) {
    if (*allele + c->bg > c->nb) return -1;
    const unsigned num = (unsigned)c->g->n_rules[head];
    unsigned raw = 0;
    for (int i = 0; i < c->bg; ++i)
        if (genome[(*allele)++]) raw += (1u << i);
    return c->first_rule[head] + (int)(raw % num);
}

static float decode_real(const ge_ctx* c, const uint8_t* genome, int* allele) {
    if (*allele + c->bc > c->nb) return 0.0f;
    unsigned long raw = 0;
    for (int i = 0; i < c->bc; ++i)
        if (genome[(*allele)++]) raw += (1ul << i);
    /* interval[0] + float(raw) * (interval[1] - interval[0]) / ((1UL << bc) - 1.0): the product is
       a float, the division and the sum double, the result returned as float */
    const float prod = (float)raw * (c->hi - c->lo);
    return (float)((double)c->lo + (double)prod / ((double)(1ul << c->bc) - 1.0));
}

static int decode_rec(const ge_ctx* c, const uint8_t* genome, int* allele, int32_t* ph, float* eph, int pos,
                      int head) {
    const int r = decode_rule(c, genome, allele, head);
    if (r < 0 || pos >= c->max_size) return 0;
    const int len = c->g->rule_len[r];
    const int32_t* syms = c->g->rule_symbols + (size_t)r * c->g->max_rule;
    for (int i = 0; i < len; ++i) {
        const int32_t s = syms[i];
        if (s >= 10000) {
            if (pos >= c->max_size) return 0; /* see the header: defined as a kill */
            ph[pos] = s;
            eph[pos] = 0.0f;
            if (c->g->t_const >= 0 && s == c->g->t_const) {
                eph[pos] = decode_real(c, genome, allele);
            } else if (s >= 20000) {
                ph[pos] = c->g->t_attribute;
                eph[pos] = (float)(s - 20000);
            }
            ++pos;
        } else {
            pos = decode_rec(c, genome, allele, ph, eph, pos, s);
            if (!pos) return 0;
        }
    }
    return pos;
}

static int ctx_init(ge_ctx* c, const ppi_ge_grammar* g, const ppi_ge_params* p) {
    int32_t bg, mx;
    if (oracle_ge_effective(g, p, &bg, &mx)) return -1;
    c->g = g;
    c->first_rule = (int32_t*)malloc(sizeof(int32_t) * (size_t)g->n_nonterminals);
    int acc = 0;
    for (int i = 0; i < g->n_nonterminals; ++i) { c->first_rule[i] = acc; acc += g->n_rules[i]; }
    c->nb = p->number_of_bits;
    c->bg = bg;
    c->bc = p->bits_per_constant;
    c->max_size = mx;
    c->lo = p->min_constant > p->max_constant ? p->max_constant : p->min_constant;
    c->hi = p->min_constant > p->max_constant ? p->min_constant : p->max_constant;
    return 0;
}

int oracle_ge_decode(const ppi_ge_grammar* g, const ppi_ge_params* p, const uint8_t* genomes, int32_t n,
                     int32_t* phenotype, float* ephemeral, int32_t* size) {
    ge_ctx c;
    if (ctx_init(&c, g, p)) return -1;
    for (int32_t i = 0; i < n; ++i) {
        int allele = 0;
        int32_t* ph = phenotype + (size_t)i * c.max_size;
        float* eph = ephemeral + (size_t)i * c.max_size;
        memset(ph, 0, sizeof(int32_t) * (size_t)c.max_size);
        memset(eph, 0, sizeof(float) * (size_t)c.max_size);
        size[i] = decode_rec(&c, genomes + (size_t)i * c.nb, &allele, ph, eph, 0, 0);
    }
    free(c.first_rule);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* breeding (ppi.cc:422-452, 680-814) and the generation loop (ppi.cc:823-945) */
/* ------------------------------------------------------------------------- */

typedef struct ge_run {
    ge_ctx c;
    const ppi_ge_params* p;
    int P, T;
    oracle_rng* rng;
} ge_run;

/* (int)(random_number() * n); a draw that rounds to 1.0 (Int() within 2^10 of 2^64, probability
   2^-54) would index one past the end in the reference and is clamped to n - 1. */
static int draw_index(oracle_rng* r, int n) {
    const int k = (int)(oracle_rng_real(r) * n);
    return k < n ? k : n - 1;
}

static int tournament(ge_run* R, oracle_rng* r, const float* fitness) {
    int w = draw_index(r, R->P);
    float fw = fitness[w];
    for (int t = 1; t < R->p->tournament_size; ++t) {
        const int k = draw_index(r, R->P);
        const float fk = fitness[k];
        if (fk < fw) { fw = fk; w = k; }
    }
    return w;
}

static void crossover(ge_run* R, oracle_rng* r, const uint8_t* f, const uint8_t* m, uint8_t* o1, uint8_t* o2) {
    const int nb = R->c.nb, bg = R->c.bg;
    if (oracle_rng_probability(r, 0.66)) {
        int a = ((int)(oracle_rng_real(r) * nb)) / bg * bg;
        int b = ((int)(oracle_rng_real(r) * nb)) / bg * bg;
        if (a > b) { const int t = a; a = b; b = t; }
        for (int i = 0; i < a; ++i) { o1[i] = f[i]; o2[i] = m[i]; }
        for (int i = a; i < b; ++i) { o1[i] = m[i]; o2[i] = f[i]; }
        for (int i = b; i < nb; ++i) { o1[i] = f[i]; o2[i] = m[i]; }
    } else {
        const int pt = (int)(oracle_rng_real(r) * nb);
        for (int i = 0; i < pt; ++i) { o1[i] = f[i]; o2[i] = m[i]; }
        for (int i = pt; i < nb; ++i) { o1[i] = m[i]; o2[i] = f[i]; }
    }
}

static void mutation(ge_run* R, oracle_rng* r, uint8_t* genome) {
    const int nb = R->c.nb, bg = R->c.bg;
    /* ceil(mutation_rate * number_of_bits) in float (a float times an int) */
    const int max_bits = (int)ceilf(R->p->mutation_rate * (float)nb);
    int num = (int)(oracle_rng_real(r) * (max_bits + 1));
    if (num > max_bits) num = max_bits;  /* a draw that rounds to 1.0 */
    if (num == 0) return;
    if (oracle_rng_probability(r, 0.75)) {
        while (num-- > 0) {
            const int bit = draw_index(r, nb);
            genome[bit] = !genome[bit];
        }
    } else {
        if (oracle_rng_probability(r, 0.10)) num = (int)(oracle_rng_real(r) * nb);
        const int shrink = ((num + (bg - 1)) / bg) * bg;
        const int start = ((int)(oracle_rng_real(r) * nb)) / bg * bg;
        const int end = start + shrink < nb ? start + shrink : nb;
        memmove(genome + start, genome + end, (size_t)(nb - end));
    }
}

/* OpenMP static schedule: thread t's iterations [*lo, *hi) of n. */
static void static_range(int n, int T, int t, int* lo, int* hi) {
    int q = n / T, rr = n % T;
    if (t < rr) { rr = 0; ++q; }
    *lo = q * t + rr;
    *hi = *lo + q;
}

static void breed(ge_run* R, const uint8_t* ants, const float* fit, uint8_t* desc, int nimm) {
    const int P = R->P, nb = R->c.nb;
    const int n_it = (P - nimm + 1) / 2;
    for (int t = 0; t < R->T; ++t) {
        int lo, hi;
        static_range(n_it, R->T, t, &lo, &hi);
        oracle_rng* r = &R->rng[t];
        for (int it = lo; it < hi; ++it) {
            const int i = nimm + 2 * it;
            const int fa = tournament(R, r, fit);
            const int mo = tournament(R, r, fit);
            uint8_t* c1 = desc + (size_t)i * nb;
            uint8_t* c2 = i < P - 1 ? desc + (size_t)(i + 1) * nb : c1;
            if (oracle_rng_real(r) < R->p->crossover_rate) {
                crossover(R, r, ants + (size_t)fa * nb, ants + (size_t)mo * nb, c1, c2);
            } else {
                memcpy(c1, ants + (size_t)fa * nb, (size_t)nb);
                if (i < P - 1) memcpy(c2, ants + (size_t)mo * nb, (size_t)nb);
            }
            mutation(R, r, c1);
            if (i < P - 1) mutation(R, r, c2);
        }
    }
}

static int evaluate(ge_run* R, const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable, const uint8_t* genomes,
                    const float* data, int32_t nlin, int32_t ncol, float* fitness, int32_t* best) {
    const int P = R->P, mx = R->c.max_size;
    int32_t* ph = (int32_t*)malloc(sizeof(int32_t) * (size_t)P * mx);
    float* eph = (float*)malloc(sizeof(float) * (size_t)P * mx);
    int32_t* sz = (int32_t*)malloc(sizeof(int32_t) * (size_t)P);
    for (int i = 0; i < P; ++i) {
        int allele = 0;
        memset(ph + (size_t)i * mx, 0, sizeof(int32_t) * (size_t)mx);
        memset(eph + (size_t)i * mx, 0, sizeof(float) * (size_t)mx);
        sz[i] = decode_rec(&R->c, genomes + (size_t)i * R->c.nb, &allele, ph + (size_t)i * mx, eph + (size_t)i * mx,
                           0, 0);
    }
    int32_t bs = 1;
    const int rc = oracle_interpret(cfg, table, ntable, ph, eph, sz, mx, P, data, nlin, ncol, R->p->alpha, fitness,
                                    best, &bs, 0, 0);
    free(ph); free(eph); free(sz);
    return rc;
}

int oracle_ge_run(const oracle_cfg* cfg, const ppi_symbol* table, int32_t ntable, const ppi_ge_grammar* g,
                  const ppi_ge_params* p, const float* data, int32_t nlin, int32_t ncol, int32_t generations,
                  uint8_t* genomes, float* fitness, float* best_history, uint8_t* best_genome) {
    ge_run R;
    if (!p || p->population_size <= 0 || p->threads <= 0 || p->tournament_size <= 0) return -1;
    if (ctx_init(&R.c, g, p)) return -1;
    R.p = p;
    R.P = p->population_size;
    R.T = p->threads;
    R.rng = (oracle_rng*)malloc(sizeof(oracle_rng) * (size_t)R.T);
    for (int t = 0; t < R.T; ++t) {
        const uint64_t s = (uint64_t)p->seed ^ (uint64_t)t;
        if (s == 0) { free(R.rng); free(R.c.first_rule); return -1; }
        oracle_rng_seed(&R.rng[t], s);
    }
    const int P = R.P, nb = R.c.nb;
    uint8_t* ants = (uint8_t*)calloc((size_t)P * nb, 1);
    uint8_t* desc = (uint8_t*)calloc((size_t)P * nb, 1);
    float* fa = (float*)malloc(sizeof(float) * (size_t)P);
    float* fd = (float*)malloc(sizeof(float) * (size_t)P);
    uint8_t* best_g = (uint8_t*)calloc((size_t)nb, 1);
    float best_f = FLT_MAX;
    int rc = 0;
    /* ppi_generate_population: genome bits (random_number() < 0.5), static over individuals */
    for (int t = 0; t < R.T; ++t) {
        int lo, hi;
        static_range(P, R.T, t, &lo, &hi);
        for (int i = lo; i < hi; ++i)
            for (int j = 0; j < nb; ++j) ants[(size_t)i * nb + j] = oracle_rng_real(&R.rng[t]) < 0.5 ? 1 : 0;
    }
    int32_t bi;
    rc = evaluate(&R, cfg, table, ntable, ants, data, nlin, ncol, fa, &bi);
    if (!rc && fa[bi] < best_f) { best_f = fa[bi]; memcpy(best_g, ants + (size_t)bi * nb, (size_t)nb); }
    if (best_history) best_history[0] = best_f;
    for (int gen = 1; gen <= generations && !rc; ++gen) {
        int nimm = 0;  /* no islands: ppi_receive_individual returns 0 */
        if (p->elitism) {
            memcpy(desc + (size_t)nimm * nb, best_g, (size_t)nb);
            fd[nimm] = best_f;
            ++nimm;
        }
        breed(&R, ants, fa, desc, nimm);
        rc = evaluate(&R, cfg, table, ntable, desc, data, nlin, ncol, fd, &bi);
        if (!rc && fd[bi] < best_f) { best_f = fd[bi]; memcpy(best_g, desc + (size_t)bi * nb, (size_t)nb); }
        if (best_history) best_history[gen] = best_f;
        uint8_t* tg = ants; ants = desc; desc = tg;
        float* tf = fa; fa = fd; fd = tf;
    }
    if (!rc) {
        if (genomes) memcpy(genomes, ants, (size_t)P * nb);
        if (fitness) memcpy(fitness, fa, sizeof(float) * (size_t)P);
        if (best_genome) memcpy(best_genome, best_g, (size_t)nb);
    }
    free(ants); free(desc); free(fa); free(fd); free(best_g); free(R.rng); free(R.c.first_rule);
    return rc;
}

/* One breeding step alone (the loop body of ppi_evolve, ppi.cc:896-936) for unit checks:
   parents/fitness -> children, with the T streams seeded seed ^ t. */
int oracle_ge_breed(const ppi_ge_grammar* g, const ppi_ge_params* p, const uint8_t* parents, const float* fitness,
                    uint8_t* children, int32_t n_immigrants) {
    ge_run R;
    if (ctx_init(&R.c, g, p)) return -1;
    R.p = p;
    R.P = p->population_size;
    R.T = p->threads;
    R.rng = (oracle_rng*)malloc(sizeof(oracle_rng) * (size_t)R.T);
    for (int t = 0; t < R.T; ++t) oracle_rng_seed(&R.rng[t], (uint64_t)p->seed ^ (uint64_t)t);
    breed(&R, parents, fitness, children, n_immigrants);
    free(R.rng);
    free(R.c.first_rule);
    return 0;
}
