// tape.h — compiles PPI prefix programs into the engine's device tape.
//
// A PPI program is a prefix expression evaluated by scanning it backwards with
// a stack (sequential.cc:143-164); the value is the final top of stack, i.e. the
// expression rooted at position 0.  Operators are pure, so any evaluation order
// of independent subtrees gives bit-identical results.  The compiler therefore
// rebuilds the expression tree and emits accumulator code in Sethi-Ullman order:
//
//   * the running value lives in one register (the accumulator);
//   * a terminal operand of an operator is folded into the operator's instruction
//     (acc = acc - x[3], acc = 2.5 * acc, ...), so leaves cost no instruction;
//   * when both operands are computed, the first one is parked in a register
//     stack slot whose index is fixed at compile time; Sethi-Ullman ordering
//     keeps the slot count at most log2(#leaves) + 1 (+1 per ternary level).
//
// Everything below the root that the reference evaluates but never reads (extra
// stack entries left under the final top) is dead and is not emitted.
#ifndef PPI_TAPE_H
#define PPI_TAPE_H

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/ppi_engine.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define PPI_HD __host__ __device__
#else
#define PPI_HD
#endif

namespace ppi {

// Instruction: 64 bits.  `code` = handler (bits 0-7) | slot (bits 8-12) |
// slot2 (bits 13-17) | ternary roles (bits 18-23); `arg` = float bits of a
// constant operand, or an attribute column.
struct Ins {
    uint32_t code;
    uint32_t arg;
};

// Device binary operators.  R* variants take the folded/parked operand as the
// LEFT operand (the accumulator is the right one).
enum BinOp : uint32_t {
    B_ADD = 0, B_SUB, B_RSUB, B_MULT, B_DIV, B_RDIV, B_MEAN, B_MAX, B_MIN, B_MOD, B_RMOD,
    B_POW, B_RPOW, B_AND, B_OR, B_XOR, B_GT, B_GE, B_LT, B_LE, B_EQ, B_NE, B_RMAX, B_RMIN,
    B_COUNT
};

// Device unary operators.
enum UnOp : uint32_t {
    U_NOT = 0, U_ABS, U_POW2, U_POW3, U_POW4, U_POW5, U_NEG, U_ROUND, U_CEIL, U_FLOOR,
    U_EXP, U_EXP10, U_EXP2, U_SIN, U_COS, U_STEP, U_SIGN, U_LOGISTIC, U_GAMMA, U_GAUSS,
    U_SQRT, U_LOG, U_LOG10, U_LOG2, U_TAN,
    U_COUNT
};

// Handlers (bits 0-7 of Ins::code): bits 5-7 are the category, bits 0-4 the
// operator, so a kernel decodes both with a shift and a mask.
enum Handler : uint32_t {
    H_LOAD_C = 0,   // acc = c
    H_LOAD_A = 1,   // acc = x[col]
    H_PUSH_C = 2,   // r[slot] = acc; acc = c
    H_PUSH_A = 3,   // r[slot] = acc; acc = x[col]
    H_TERN = 4,     // acc = select(cond, then, else) over {acc, r[slot], r[slot2]}
    H_END = 5,      // end of program: code[code_off + ncode] of every program
    H_UN = 32,      // + UnOp:        acc = f(acc)
    H_BIN_C = 64,   // + BinOp:       acc = f(acc, c)
    H_BIN_A = 96,   // + BinOp:       acc = f(acc, x[col])
    H_BIN_S = 128,  // + BinOp:       acc = f(acc, r[slot])
    H_LIMIT = 160
};
static_assert(U_COUNT <= 32 && B_COUNT <= 32, "operators fit the 5-bit operator field");

constexpr int kMaxSlots = 16;

// Operand count of an engine opcode, and the device operator of a unary / binary opcode (shared by
// the host compiler and the device compiler of the grammatical-evolution path).
inline PPI_HD int arity(int32_t op) {
    if (op == PPI_OP_IF_THEN_ELSE) return 3;
    if (op >= PPI_OP_AND && op <= PPI_OP_DIV) return 2;
    if (op >= PPI_OP_NOT && op <= PPI_OP_TAN) return 1;
    return 0;
}

inline PPI_HD uint32_t unop_of(int32_t op) {
    switch (op) {
    case PPI_OP_NOT: return U_NOT;     case PPI_OP_ABS: return U_ABS;
    case PPI_OP_POW2: return U_POW2;   case PPI_OP_POW3: return U_POW3;
    case PPI_OP_POW4: return U_POW4;   case PPI_OP_POW5: return U_POW5;
    case PPI_OP_NEG: return U_NEG;     case PPI_OP_ROUND: return U_ROUND;
    case PPI_OP_CEIL: return U_CEIL;   case PPI_OP_FLOOR: return U_FLOOR;
    case PPI_OP_EXP: return U_EXP;     case PPI_OP_EXP10: return U_EXP10;
    case PPI_OP_EXP2: return U_EXP2;   case PPI_OP_SIN: return U_SIN;
    case PPI_OP_COS: return U_COS;     case PPI_OP_STEP: return U_STEP;
    case PPI_OP_SIGN: return U_SIGN;   case PPI_OP_LOGISTIC: return U_LOGISTIC;
    case PPI_OP_GAMMA: return U_GAMMA; case PPI_OP_GAUSS: return U_GAUSS;
    case PPI_OP_SQRT: return U_SQRT;   case PPI_OP_LOG: return U_LOG;
    case PPI_OP_LOG10: return U_LOG10; case PPI_OP_LOG2: return U_LOG2;
    default: return U_TAN;
    }
}

// Binary operator with the accumulator as the left operand (rev=false) or as the
// right operand (rev=true); the other operand is the folded/parked value.
inline PPI_HD uint32_t binop_of(int32_t op, bool rev) {
    switch (op) {
    case PPI_OP_AND: return B_AND;
    case PPI_OP_OR: return B_OR;
    case PPI_OP_XOR: return B_XOR;
    case PPI_OP_GREATER: return rev ? B_LT : B_GT;
    case PPI_OP_GREATEREQUAL: return rev ? B_LE : B_GE;
    case PPI_OP_LESS: return rev ? B_GT : B_LT;
    case PPI_OP_LESSEQUAL: return rev ? B_GE : B_LE;
    case PPI_OP_EQUAL: return B_EQ;
    case PPI_OP_NOTEQUAL: return B_NE;
    case PPI_OP_ADD: return B_ADD;
    case PPI_OP_SUB: return rev ? B_RSUB : B_SUB;
    case PPI_OP_MULT: return B_MULT;
    case PPI_OP_MEAN: return B_MEAN;
    case PPI_OP_MAX: return rev ? B_RMAX : B_MAX;
    case PPI_OP_MIN: return rev ? B_RMIN : B_MIN;
    case PPI_OP_MOD: return rev ? B_RMOD : B_MOD;
    case PPI_OP_POW: return rev ? B_RPOW : B_POW;
    default: return rev ? B_RDIV : B_DIV;
    }
}


// Per-program header on the device (16 bytes).
struct ProgMeta {
    uint32_t code_off;  // first instruction in the tape
    uint32_t ncode;     // instructions, excluding the H_END that follows them
    int32_t size;       // original phenotype size (for alpha*size and GPop accounting)
    uint32_t flags;     // kEmpty | kInvalid; bits 8-15: slots used
};
enum : uint32_t { kEmpty = 1u, kInvalid = 2u };

struct SymbolTable {
    // symbol -> ppi_symbol, for symbols in [10000, 10000 + 8192) by direct index,
    // others through a sorted list.
    std::vector<int32_t> direct;  // index+1 into entries, 0 = none
    std::vector<ppi_symbol> entries;
    std::vector<std::pair<int32_t, int32_t>> other;  // (symbol, index) sorted
    void build(const ppi_symbol* t, int32_t n);
    const ppi_symbol* find(int32_t symbol) const;
};

struct CompileStats {
    int64_t total_size = 0;     // sum of phenotype sizes
    int64_t total_code = 0;     // sum of tape instructions
    int32_t max_slots = 0;
    int32_t max_code = 0;
    int32_t n_invalid = 0;
    int32_t n_empty = 0;
    int32_t max_col = -1;       // largest input column referenced (derived columns excluded)
    int64_t derived_uses = 0;   // unary-of-column nodes read from derived columns
    int64_t folded_nodes = 0;   // operator nodes inside folded constant subtrees
};

// Chains of unary operators over input columns, precomputed once per dataset:
// the compiler turns a unary node over an attribute leaf into a leaf that reads
// derived column base + j, which holds op(x[src]) for every row, and a unary
// node over such a leaf into another derived column (sin(cos(x3)) reads one
// column).  The values are computed on the device by the same operator code, in
// column order (a column's source is an input column or an earlier derived
// column), so they are bit-identical to evaluating the nodes.
//
// A chain is named by its key: bits 0-15 = input column + 1, then 5 bits per
// operator, innermost first ((unary op + 1) << (16 + 5 i)), at most
// kChainMaxOps operators.  plan_derived() counts the chains of a population in
// parallel and gives columns to the most frequent chains not yet in the table
// (ties: fewer operators, then smaller key; a chain's prefix always ranks
// first), so the table is deterministic and, while compiling, read-only.
constexpr int kChainMaxOps = 9;

// Open-addressing hash map from non-zero 64-bit keys (chain keys always carry a column + 1 in
// their low bits) to V: linear probing, power-of-two capacity at most half full.  The chain
// tables are looked up once per unary node over a column while compiling and counted once per
// such node while planning; a node-based std::unordered_map made that the compiler's largest cost.
template <typename V>
struct FlatMap64 {
    std::vector<uint64_t> keys;  // 0 = empty slot
    std::vector<V> vals;
    size_t count = 0;
    int shift = 64;

    size_t slot_of(uint64_t k) const { return (size_t)((k * 0x9E3779B97F4A7C15ull) >> shift); }
    const V* find(uint64_t k) const {
        if (keys.empty()) return nullptr;
        const size_t mask = keys.size() - 1;
        for (size_t i = slot_of(k);; i = (i + 1) & mask) {
            if (keys[i] == k) return &vals[i];
            if (keys[i] == 0) return nullptr;
        }
    }
    bool contains(uint64_t k) const { return find(k) != nullptr; }
    // The value of key k, inserted as V{} if absent.
    V& at(uint64_t k) {
        if ((count + 1) * 2 > keys.size()) rehash(keys.empty() ? 1024 : keys.size() * 2);
        const size_t mask = keys.size() - 1;
        size_t i = slot_of(k);
        for (; keys[i] != 0 && keys[i] != k; i = (i + 1) & mask) {}
        if (keys[i] == 0) {
            keys[i] = k;
            vals[i] = V{};
            ++count;
        }
        return vals[i];
    }
    void rehash(size_t cap) {
        std::vector<uint64_t> ok;
        std::vector<V> ov;
        ok.swap(keys);
        ov.swap(vals);
        keys.assign(cap, 0);
        vals.assign(cap, V{});
        shift = 64 - __builtin_ctzll((unsigned long long)cap);
        count = 0;
        for (size_t j = 0; j < ok.size(); ++j)
            if (ok[j]) at(ok[j]) = ov[j];
    }
    void clear() {
        keys.clear();
        vals.clear();
        count = 0;
        shift = 64;
    }
    size_t size() const { return count; }
    template <class F>
    void for_each(F&& f) const {
        for (size_t j = 0; j < keys.size(); ++j)
            if (keys[j]) f(keys[j], vals[j]);
    }
    bool operator==(const FlatMap64& o) const {
        if (count != o.count) return false;
        bool same = true;
        for_each([&](uint64_t k, const V& v) {
            const V* w = o.find(k);
            same = same && w && *w == v;
        });
        return same;
    }
};

struct DerivedColumns {
    int32_t base = 0;  // first derived column (the dataset's column count)
    int32_t cap = 0;   // most derived columns (0: off)
    uint32_t epoch = 0;        // counts resets: columns staged under an older epoch are stale
    uint32_t staged_epoch = 0; // epoch of the columns on the device (stage_derived_columns)
    bool refill = true;        // false: never reset by plan_derived (compiled tapes of earlier
                               // populations still read the columns: the batched evaluator)
    struct Col { uint32_t op; int32_t src; };
    std::vector<Col> cols;                          // column base + j = op(column src), src < base + j
    FlatMap64<int32_t> index;                       // chain key -> column
    void reset(int32_t b, int32_t c) {
        base = b;
        cap = c;
        cols.clear();
        index.clear();
        ++epoch;
    }
    // Column holding the chain, or -1.
    int32_t find(uint64_t key) const {
        const int32_t* c = index.find(key);
        return c ? *c : -1;
    }
};

// Adds columns for the unary chains of programs [0, n) to `derived` (see above);
// `threads` workers count the chains.  A full table whose columns cover less than
// kDerivedKeepShare of the population's chain occurrences is reset (its epoch
// advances) and refilled from this population: over a long run the table follows
// the chains the programs use now instead of the first generations'.
constexpr double kDerivedKeepShare = 0.5;
void plan_derived(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                  int32_t stride, int32_t n, int32_t ncol, int threads, DerivedColumns* derived);

// Constant subtrees (every leaf a constant, at least one operator) folded out of
// the programs: each is a program of its own in `meta`/`code` (same format, ending
// in H_END), and the value it evaluates to is the argument of instruction at[j]
// of the main tape, which the compiler leaves as a placeholder.  The folds are
// evaluated once per population by the engine's own interpreter code on the
// device (fold_kernel), so the folded tape computes the same bits as the
// unfolded one; apply_folds() writes the values in.
struct FoldSet {
    std::vector<ProgMeta> meta;
    std::vector<Ins> code;
    std::vector<uint32_t> at;
};

struct Tape {
    std::vector<ProgMeta> meta;
    std::vector<Ins> code;
    std::vector<uint32_t> const_pos;  // per instruction: phenotype position + 1 of a T_CONST operand, else 0
    CompileStats stats;
    FoldSet folds;
};

// Writes the values of the tape's folds (value[j] of fold j) into their
// placeholder instructions.
void apply_folds(Tape* t, const float* value);

// Compiles programs [0, n) into `out` (appending).  ncol bounds attribute columns
// (an out-of-range column evaluates as NaN).  With `derived`, the population's
// unary chains are planned first (plan_derived) and read from derived columns.
// With `fold`, constant subtrees go to out->folds (not with want_const_pos: the
// adjoint needs every T_CONST in the tape).  Returns false with `err` set if a
// program needs more than kMaxSlots register slots.
bool compile_population(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral,
                        const int32_t* size, int32_t stride, int32_t n, int32_t ncol, bool want_const_pos,
                        Tape* out, std::string* err, DerivedColumns* derived = nullptr, bool fold = false);

// As compile_population with the derived columns already planned: `derived` is only read, so
// several populations sharing it may be compiled at once (one thread each).
bool compile_population_planned(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral,
                                const int32_t* size, int32_t stride, int32_t n, int32_t ncol, Tape* out,
                                std::string* err, const DerivedColumns* derived, bool fold);

// Multi-threaded front end: plans the derived columns and compiles with `threads` workers.
bool compile_population_mt(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral,
                           const int32_t* size, int32_t stride, int32_t n, int32_t ncol, bool want_const_pos,
                           int threads, Tape* out, std::string* err, DerivedColumns* derived = nullptr,
                           bool fold = false);

// ---------------------------------------------------------------------------
// Reverse-mode adjoint tape: d(fitness)/d(ephemeral T_CONST values).
//
// Every T_CONST leaf is its own parameter and a tree node has one parent, so the
// derivative of the program value with respect to a constant is the product of
// the local partials on the constant's path to the root.  Paths through
// operators whose derivative is zero everywhere it exists (logic, comparisons,
// ROUND/CEIL/FLOOR/STEP/SIGN, the condition of IF_THEN_ELSE) are cut: those
// constants have gradient 0 and are left out.
//
// The forward tape is the ordinary accumulator tape.  Stored partials are
// numbered e = 0, 1, ... in the order the reverse sweep reads them; each is the
// derivative of one instruction with respect to its accumulator operand or its
// other operand (for IF_THEN_ELSE the then- and else-selectors).  Partials that
// are constants (ADD, SUB, MEAN, NEG, one side of MOD) are not stored.  Each
// segment lists, in instruction order, the instructions that store one of its
// partials and the slots they store to, so the forward pass runs plain steps
// between them and checks nothing per instruction.
//
// The reverse sweep walks the active tree root-first.  Adjoints live in
// register slots with stack discipline (a node's last child reuses its slot), and
// a constant's contribution adj(parent) * partial is summed into the lane's
// gradient accumulator k.  When a kernel keeps B partials per row, partials e
// with e / B == s belong to segment s: the kernel re-runs the forward pass (up
// to fwd_end) once per segment, keeping only that segment's partials, and then
// runs the segment's reverse instructions.  Adjoint slots persist across
// segments.  A program with more than kAdjMaxConsts live constants is split
// into several jobs, each owning up to kAdjMaxConsts constants.
// ---------------------------------------------------------------------------
constexpr int kAdjMaxConsts = 16;  // gradient accumulators per lane
constexpr int kAdjSlots = 12;      // adjoint register slots per row

enum : uint32_t { kAdjFitness = 1u << 4 };  // AdjJob::flags: this job also accumulates the fitness

struct AdjJob {
    uint32_t code_off;    // forward tape of the program (as ProgMeta::code_off)
    uint32_t ncode;
    uint32_t seg_off;     // segments of this job: segs[seg_off + s], s < nseg
    uint32_t nseg;
    int32_t prog;
    uint32_t const_base;  // gradient index of this job's constant 0
    uint32_t nconst;
    uint32_t flags;       // kEmpty / kInvalid of the program, kAdjFitness
    uint32_t cost;        // ncode * nseg + reverse instructions (scheduling weight)
    uint32_t cols;        // bit c: the program reads input column c (c < 32)
    uint32_t pad_[2];
};

struct AdjSeg {
    uint32_t rev_begin, rev_end;  // reverse instructions [rev_begin, rev_end) (absolute)
    uint32_t fwd_end;             // forward instructions [0, fwd_end) needed by the segment
    uint32_t st_begin, st_end;    // stores [st_begin, st_end) (absolute), increasing pc
    uint32_t pad_[3];
};

// A forward instruction that stores partials of its segment: pc relative to the
// program; slots = (slot of the accumulator-side partial + 1) | (slot of the
// other-side partial + 1) << 16, 0 = that side is not stored in the segment.
struct AdjStore {
    uint32_t pc;
    uint32_t slots;
};

// Reverse instruction.  w: bits 0-1 kind (kRevAdj: adj[dst] = adj[src] * p;
// kRevGrad: g[dst] += adj[src] * p), bit 2 = p is stored partial e (else the
// float bits in `p`), bits 4-7 src slot, bits 8-12 dst slot / constant,
// bits 16-31 e.
struct RevIns {
    uint32_t w;
    uint32_t p;
};
enum : uint32_t { kRevAdj = 0u, kRevGrad = 1u, kRevStored = 4u };

struct AdjConst {
    int32_t prog;   // program
    int32_t pos;    // phenotype position of the T_CONST
    uint32_t ins;   // tape instruction whose argument holds the constant
    uint32_t pad_;
};

struct AdjTape {
    Tape fwd;                          // forward tape (identical to compile_population's), with const_pos
    std::vector<AdjJob> jobs;          // program order; a program's jobs are consecutive
    std::vector<AdjStore> stores;
    std::vector<AdjSeg> segs;
    std::vector<RevIns> rev;
    std::vector<AdjConst> consts;      // gradient index -> constant
    std::vector<uint32_t> prog_jobs;   // n + 1 offsets into jobs
    std::vector<uint32_t> prog_consts; // n + 1 offsets into consts
    int32_t max_seg = 0;               // most segments of one job
    int32_t max_adj_slots = 0;
    int64_t stored_partials = 0;       // sum over jobs of stored partials
    int64_t total_rev = 0;
};

// Compiles programs [0, n) for the adjoint kernels, whose partial store holds B
// (a power of two) partials per row.  Returns false with `err` set on a program
// the forward compiler rejects or that needs more than kAdjSlots adjoint slots.
bool compile_adjoint(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                     int32_t stride, int32_t n, int32_t ncol, int32_t B, AdjTape* out, std::string* err);

}  // namespace ppi

#endif
