// Test-only host interpreter of the engine's compiled tape (no GPU involved).
// Compiling the prefix programs with ppi::compile_population and executing the
// tape here must give, bit for bit, the oracle's prefix-stack values: this pins
// the compiler's tree rebuild, Sethi-Ullman ordering, operand orientation and
// register-slot assignment on the CPU.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../ppi-auto_amd/csrc/tape.h"

using namespace ppi;

static float bf(bool b) { return b ? 1.0f : 0.0f; }
static float dv(float x, float y, int P) { return (P && y == 0.0f) ? 1.0f : x / y; }
static float sq(float x, int P) { return (P && x < 0.0f) ? 1.0f : (float)std::sqrt((double)x); }
static float lg(float x, int P, int k) {
    if (P && x < 1.0f) return 1.0f;
    return k == 0 ? (float)std::log((double)x) : (k == 1 ? (float)std::log10((double)x) : (float)std::log2((double)x));
}

static float un(uint32_t u, float a, int P) {
    const float PI_2 = 1.57079637050629f, PI = 3.14159274101257f, E = 2.71828174591064f;
    switch (u) {
    case U_NOT: return bf(!(a != 0.0f));
    case U_ABS: return std::fabs(a);
    case U_POW2: return a * a;
    case U_POW3: return a * a * a;
    case U_POW4: return a * a * a * a;
    case U_POW5: return a * a * a * a * a;
    case U_NEG: return -a;
    case U_ROUND: return std::round(a);
    case U_CEIL: return std::ceil(a);
    case U_FLOOR: return std::floor(a);
    case U_EXP: return (float)std::exp((double)a);
    case U_EXP10: return (float)::exp10((double)a);
    case U_EXP2: return (float)std::exp2((double)a);
    case U_SIN: return (float)std::sin((double)a);
    case U_COS: return (float)std::cos((double)a);
    case U_STEP: return bf(a >= 0.0f);
    case U_SIGN: return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f);
    case U_LOGISTIC: return 1.0f / (1.0f + (float)std::exp((double)-a));
    case U_GAMMA: {
        float base = (a / E) * sq((float)((double)a * std::sinh((double)(1.0f / a))), P);
        return (float)(std::pow((double)base, (double)a) * (double)sq(2.0f * PI / a, P));
    }
    case U_GAUSS: return (float)std::exp((double)(-a * a));
    case U_SQRT: return sq(a, P);
    case U_LOG: return lg(a, P, 0);
    case U_LOG10: return lg(a, P, 1);
    case U_LOG2: return lg(a, P, 2);
    default:
        if (P && (std::fabs(a) == PI_2 || std::fabs(a) == 3.0f * PI_2)) return 1.0f;
        return (float)std::tan((double)a);
    }
}

static float bin(uint32_t op, float a, float b, int P) {
    switch (op) {
    case B_ADD: return a + b;
    case B_SUB: return a - b;
    case B_RSUB: return b - a;
    case B_MULT: return a * b;
    case B_DIV: return dv(a, b, P);
    case B_RDIV: return dv(b, a, P);
    case B_MEAN: return (a + b) / 2.0f;
    case B_MAX: return (float)std::fmax((double)a, (double)b);
    case B_RMAX: return (float)std::fmax((double)b, (double)a);
    case B_MIN: return (float)std::fmin((double)a, (double)b);
    case B_RMIN: return (float)std::fmin((double)b, (double)a);
    case B_MOD: return (float)std::fmod((double)a, (double)b);
    case B_RMOD: return (float)std::fmod((double)b, (double)a);
    case B_POW: return (float)std::pow((double)a, (double)b);
    case B_RPOW: return (float)std::pow((double)b, (double)a);
    case B_AND: return bf(a != 0.0f && b != 0.0f);
    case B_OR: return bf(a != 0.0f || b != 0.0f);
    case B_XOR: return bf((a != 0.0f) != (b != 0.0f));
    case B_GT: return bf(a > b);
    case B_GE: return bf(a >= b);
    case B_LT: return bf(a < b);
    case B_LE: return bf(a <= b);
    case B_EQ: return bf(a == b);
    default: return bf(a != b);
    }
}

// One program of the tape on one row (row may be NULL for programs without attributes).
static float run_program(const std::vector<Ins>& code, const ProgMeta& m, const float* row, int P) {
    float acc = NAN, reg[kMaxSlots];
    for (int k = 0; k < kMaxSlots; ++k) reg[k] = NAN;
    for (uint32_t pc = 0; pc < m.ncode; ++pc) {
        const Ins in = code[m.code_off + pc];
        const uint32_t h = in.code & 0xffu, s1 = (in.code >> 8) & 31u, s2 = (in.code >> 13) & 31u;
        float cv;
        std::memcpy(&cv, &in.arg, 4);
        auto term = [&](bool attr) { return attr ? row[in.arg] : cv; };
        if (h >= H_BIN_S) acc = bin(h - H_BIN_S, acc, reg[s1], P);
        else if (h >= H_BIN_A) acc = bin(h - H_BIN_A, acc, term(true), P);
        else if (h >= H_BIN_C) acc = bin(h - H_BIN_C, acc, term(false), P);
        else if (h >= H_UN) acc = un(h - H_UN, acc, P);
        else if (h == H_LOAD_C || h == H_LOAD_A) acc = term(h == H_LOAD_A);
        else if (h == H_PUSH_C || h == H_PUSH_A) { reg[s1] = acc; acc = term(h == H_PUSH_A); }
        else {
            const float v[3] = {acc, reg[s1], reg[s2]};
            const float c = v[(in.code >> 18) & 3u], th = v[(in.code >> 20) & 3u], el = v[(in.code >> 22) & 3u];
            acc = (c != 0.0f) ? th : el;
        }
    }
    return acc;
}

// derive bit 0: compile with derived columns (chains of unary operators over input
// columns read precomputed columns), computed here on the host from the same
// operator code, in column order.  derive bit 1: fold constant subtrees; the fold
// programs run here first and their values are written into the tape, as the
// engine does on the device.  derived_out (may be NULL) receives the number of
// derived columns, the number of them whose source is another derived column, the
// number of folded subtrees and the operator nodes inside them.
extern "C" int tape_eval_ex(const ppi_symbol* table, int ntable, const int32_t* ph, const float* eph, const int32_t* size,
                            int stride, int n, const float* data, int nlin, int ncol, int P, float* out,
                            int32_t* slots, int32_t* ncode, int derive, int32_t* derived_out) {
    SymbolTable st;
    st.build(table, ntable);
    Tape t;
    std::string err;
    DerivedColumns dc;
    dc.reset(ncol, (derive & 1) ? 4096 : 0);
    if (!compile_population(st, ph, eph, size, stride, n, ncol, false, &t, &err, (derive & 1) ? &dc : nullptr,
                            (derive & 2) != 0))
        return -1;
    if (!t.folds.meta.empty()) {
        std::vector<float> value(t.folds.meta.size());
        for (size_t j = 0; j < value.size(); ++j) value[j] = run_program(t.folds.code, t.folds.meta[j], nullptr, P);
        apply_folds(&t, value.data());
    }
    if (derived_out) {
        derived_out[0] = (int32_t)dc.cols.size();
        derived_out[1] = 0;
        for (const auto& c : dc.cols) derived_out[1] += c.src >= ncol ? 1 : 0;
        derived_out[2] = (int32_t)t.folds.meta.size();
        derived_out[3] = (int32_t)t.stats.folded_nodes;
    }
    const int wide = ncol + (int)dc.cols.size();
    std::vector<float> ext((size_t)nlin * wide);
    for (int r = 0; r < nlin; ++r) {
        for (int c = 0; c < ncol; ++c) ext[(size_t)r * wide + c] = data[(size_t)r * ncol + c];
        for (size_t j = 0; j < dc.cols.size(); ++j)  // in order: a chain's source column comes first
            ext[(size_t)r * wide + ncol + j] = un(dc.cols[j].op, ext[(size_t)r * wide + dc.cols[j].src], P);
    }
    data = ext.data();
    ncol = wide;
    for (int i = 0; i < n; ++i) {
        const ProgMeta& m = t.meta[i];
        slots[i] = (int32_t)((m.flags >> 8) & 0xff);
        ncode[i] = (int32_t)m.ncode;
        for (int r = 0; r < nlin; ++r) {
            const float* row = data + (size_t)r * ncol;
            out[(size_t)i * nlin + r] = (m.flags & kEmpty) ? NAN : run_program(t.code, m, row, P);
        }
    }
    return 0;
}

extern "C" int tape_eval(const ppi_symbol* table, int ntable, const int32_t* ph, const float* eph, const int32_t* size,
                         int stride, int n, const float* data, int nlin, int ncol, int P, float* out,
                         int32_t* slots, int32_t* ncode) {
    return tape_eval_ex(table, ntable, ph, eph, size, stride, n, data, nlin, ncol, P, out, slots, ncode, 0, nullptr);
}

// The derived-column table across populations (plan_derived's refill rule): out[0] columns and
// out[1] epoch after population A; out[2] epoch after A again; out[3] epoch and out[4] columns
// after B; out[5] 1 if every chain key of the table after B occurs in B's programs as counted by a
// fresh table; out[6] epoch after A with refills turned off.
extern "C" int derived_refill(const ppi_symbol* table, int ntable, const int32_t* phA, const float* ephA,
                              const int32_t* sizeA, int strideA, int nA, const int32_t* phB, const float* ephB,
                              const int32_t* sizeB, int strideB, int nB, int ncol, int cap, int32_t* out) {
    SymbolTable st;
    st.build(table, ntable);
    DerivedColumns dc;
    dc.reset(ncol, cap);
    const uint32_t e0 = dc.epoch;
    plan_derived(st, phA, ephA, sizeA, strideA, nA, ncol, 1, &dc);
    out[0] = (int32_t)dc.cols.size();
    out[1] = (int32_t)(dc.epoch - e0);
    plan_derived(st, phA, ephA, sizeA, strideA, nA, ncol, 1, &dc);
    out[2] = (int32_t)(dc.epoch - e0);
    plan_derived(st, phB, ephB, sizeB, strideB, nB, ncol, 1, &dc);
    out[3] = (int32_t)(dc.epoch - e0);
    out[4] = (int32_t)dc.cols.size();
    DerivedColumns fresh;
    fresh.reset(ncol, cap);
    plan_derived(st, phB, ephB, sizeB, strideB, nB, ncol, 1, &fresh);
    out[5] = fresh.index == dc.index ? 1 : 0;
    dc.refill = false;
    const uint32_t e1 = dc.epoch;
    plan_derived(st, phA, ephA, sizeA, strideA, nA, ncol, 1, &dc);
    out[6] = (int32_t)(dc.epoch - e1);
    return 0;
}
