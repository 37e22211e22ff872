// ge.h — grammatical evolution on the device: shapes shared by the runtime (ge.cc) and the
// kernels (kern_ge.hip).
//
// Genomes are bit-packed: individual i's allele j is bit (j & 31) of word i * W + (j >> 5), so
// the decoder reads 32 alleles per load and crossover / shrink move words, not bytes.  The
// reference keeps one bool per allele (individual:11, GENOME_TYPE = bool); the C ABI converts.
#ifndef PPI_GE_H
#define PPI_GE_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tape.h"

namespace ppi {

// The decoder's grammar on the device (gramatica[] / tamanhos[] of the grammar header).
struct GeDevGrammar {
    const int32_t* first_rule;  // per non-terminal: its first rule
    const int32_t* n_rules;     // per non-terminal: tamanhos[]
    const int32_t* rule_len;    // per rule
    const int32_t* rule_sym;    // rule r: [r * max_rule, r * max_rule + rule_len[r])
    int32_t max_rule;
    int32_t t_const;            // -1: the grammar has no T_CONST
    int32_t t_attribute;
};

// ppi_init's options after the reference's adjustments (ppi.cc:258-307).
struct GeShape {
    int32_t P;         // population size
    int32_t nb;        // alleles per genome
    int32_t W;         // 32-bit words per genome
    int32_t bg, bc;    // bits per gene (raised to the grammar's need), bits per constant
    int32_t max_size;  // phenotype stride (max_size_phenotype)
    float lo, hi;      // constant interval
    float cp;          // crossover rate
    int32_t ts;        // tournament size
    int32_t T;         // random streams (the reference's OpenMP threads)
    int32_t max_bits;  // ceil(mutation_rate * nb) in float (ppi.cc:743)
};

// One iteration of the breeding loop (ppi.cc:896-936): parents, crossover cut, the mutation of each
// child.  xo: 0 = clone (no crossover), 1 = one point at a, 2 = two points [a, b).  mut[c]: 0 none,
// 1 = nflip[c] bit flips (flips[(it * 2 + c) * max_bits + k]), 2 = shrink [s0[c], s1[c]).
struct GePlan {
    int32_t father, mother;
    int32_t xo, a, b;
    int32_t mut[2];
    int32_t nflip[2];
    int32_t s0[2], s1[2];
};

// Per-program output of the device compiler before the tape is packed.
struct GeCompiled {
    uint32_t ncode;   // instructions, excluding the H_END
    uint32_t flags;   // kEmpty | kInvalid | slots << 8
    uint32_t end_arg; // argument of the program's H_END (as tape.cc's put_end)
    int32_t max_col;  // largest input column read (-1: none)
};

// Scratch of the device compiler, per program (strides in elements).
struct GeCompileScratch {
    uint32_t* node;    // 4 words per node: arg; ch0 | ch1 << 16; ch2 | fold << 16; op | regs << 8 | leaf_attr << 16
    int32_t* stk;      // value-stack of node indices / emit frames / fold-evaluation frames
    float* fval;       // fold-evaluation values
    Ins* stage;        // the program's instructions
    int64_t node_stride, stk_stride, fval_stride, stage_stride;
};

// Population-wide compile statistics (atomics).
struct GeCompileTotals {
    int64_t total_size;
    int64_t total_code;
    int32_t max_slots;
    int32_t max_col;
    int32_t error;      // 1: a program needs more than kMaxSlots register slots; 2: size outside [0, stride]
    int32_t pad_;
};

// The random streams in parallel.  A stream's state advances by a linear map T over GF(2)^128
// (XorShift128+'s shifts and XORs; the output addition is applied to a state, not part of it), so the
// state kDrawBlock * k draws ahead is T^(kDrawBlock * k) applied to it: `jump` holds the columns of
// T^(kDrawBlock * 2^j), j < kJumpLevels, and one thread per block of draws reaches its block by at
// most kJumpLevels matrix-vector products, then draws its block in order.  A breeding step draws a
// data-dependent number of values per iteration (the mutation draws as many positions as it drew
// bits); the iteration that starts at draw p ends at p + L(p), L computed for every p in parallel,
// and the start of each stream's l-th iteration is found by pointer doubling over p -> p + L(p).
constexpr int kDrawBlock = 256;
constexpr int kJumpLevels = 24;

struct GeDraws {
    const uint64_t* jump;   // [kJumpLevels][128][2]
    uint64_t* rng;          // [T][2] stream states, advanced past the draws a phase consumed
    double* draw;           // [T][cap] Real() of each stream's draws, in order (breeding)
    uint64_t* bstate;       // [T][nblk + 1][2] state at the start of each block
    int64_t cap;            // draws per stream, a multiple of kDrawBlock
    int32_t nblk;           // cap / kDrawBlock
    int32_t* nxt;           // [levels][T][cap + 1] start of the iteration 2^j iterations after one at p
    int32_t levels;
    int64_t* end;           // [T] draws the phase consumed
};

// Host: columns of T^(kDrawBlock * 2^j), j < kJumpLevels (kJumpLevels * 256 words).
void ge_jump_tables(uint64_t* out);
// Generation 0 (ppi_generate_population): each stream draws its individuals' alleles in order.
hipError_t ge_launch_generate(const GeShape& S, const GeDraws& D, uint32_t* genomes, hipStream_t st);
// One breeding step's plan (every draw of ppi_evolve's loop body, tournaments resolved), offspring
// built separately by ge_launch_offspring.
hipError_t ge_launch_plan(const GeShape& S, const GeDraws& D, const float* fitness, int32_t nimm, GePlan* plan,
                          int32_t* flips, hipStream_t st);
// Draws per stream a breeding step may consume at most, per iteration.
inline int64_t ge_draws_per_iteration(const GeShape& S) {
    // tournaments, crossover test + Probability(0.66) + two cuts, per child: bit count + Probability(0.75)
    // + the positions, or Probability(0.10) + a new count + the start of the shrink
    return 2 * (int64_t)S.ts + 4 + 2 * (2 + (int64_t)(S.max_bits > 3 ? S.max_bits : 3));
}
hipError_t ge_launch_offspring(const GeShape& S, const GePlan* plan, const int32_t* flips, const uint32_t* parents,
                               uint32_t* children, int32_t nimm, hipStream_t st);
hipError_t ge_launch_decode(const GeShape& S, const GeDevGrammar& G, const uint32_t* genomes, int32_t* ph,
                            float* eph, int32_t* size, int32_t* frames, int64_t frame_stride, hipStream_t st);
hipError_t ge_launch_compile(const int32_t* ph, const float* eph, const int32_t* size, int32_t stride, int32_t n,
                             const int32_t* sym_op, const float* sym_val, int32_t ncol, bool fold, bool P, bool N,
                             const GeCompileScratch& X, GeCompiled* out, GeCompileTotals* tot, hipStream_t st);
// Packs the staged programs into a tape: meta[i] = {code_off, ncode, size, flags}, instructions and
// H_END per program, kTapePad padding entries after the last.  offs: n + 1 scratch words.
hipError_t ge_launch_pack(const GeCompileScratch& X, const GeCompiled* comp, const int32_t* size, int32_t n,
                          uint32_t* offs, ProgMeta* meta, Ins* code, GeCompileTotals* tot, hipStream_t st);
// Copies `words` words from src to dst when fitness[*best] < *best_fit (best-individual tracking).
hipError_t ge_launch_track_best(const float* fitness, const int32_t* best, const uint32_t* genomes, int32_t W,
                                float* best_fit, uint32_t* best_genome, hipStream_t st);

// The reference's epilogue on the device (accelerator.cc:843-889, sequential.cc:185-208):
// vec[i] from the kernels' raw values with alpha*size, then the strategy's ranking into
// index[0..k): PickNBest over the vector (SEQ), or best_individual over work-groups of best_local
// and PickNBest over the group winners (FP: the vector; PP, PDP: the raw values).  k <= 64, already
// clipped by the caller (to n, or to the group count); group_val / group_idx hold a value per group.
constexpr int kRankMax = 64;
hipError_t launch_rank(const float* raw, const ProgMeta* meta, int32_t n, int32_t strategy, float alpha,
                       int32_t best_local, int32_t k, float* vec, float* group_val, int32_t* group_idx,
                       int32_t* index, hipStream_t st);

}  // namespace ppi

#endif
