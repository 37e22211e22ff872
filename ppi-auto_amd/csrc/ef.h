// ef.h — user error functions ERROR(X,Y) given as C expression text (definitions.h:18,
// CMakeLists.txt:26-33: the reference pastes -DEF=<expression> into its kernels and into
// sequential.cc, where X is the program's prediction and Y the row's target, both float).
//
// The expression is compiled on the host into a postfix tape of typed operations that the
// device interprets per row.  Typing follows C: integer literals are int, unsuffixed floating
// literals double, suffixed ones float; binary operators convert their operands by the usual
// arithmetic conversions (int < float < double); comparisons and && || ! give int; the
// conditional operator converts both branches to their common type; casts (int), (float),
// (double).  Functions, as sequential.cc sees them (<cmath> without `using namespace std`:
// the C library's double routines, and the C99 float ones with the f suffix): fabs sqrt exp
// log pow sin cos tan floor ceil round trunc fmin fmax take and return double; fabsf sqrtf
// floorf ceilf roundf truncf fminf fmaxf take and return float (exact operations); abs takes
// the overload of its argument's type (libstdc++'s <stdlib.h>).  The C library's float
// transcendentals (expf logf sinf cosf tanf powf) are refused: they are not correctly rounded,
// so their results cannot be reproduced from their definition.  The result is converted to
// float (the reference assigns it to `float error`).
#ifndef PPI_EF_H
#define PPI_EF_H

#include <cstdint>
#include <string>
#include <vector>

namespace ppi {

enum EfType : uint8_t { EF_INT = 0, EF_FLOAT = 1, EF_DOUBLE = 2 };

enum EfOp : uint8_t {
    EF_X, EF_Y, EF_IMM,                                  // push X, Y, imm (of `type`)
    EF_CVT,                                              // top converted to `type`
    EF_NEG, EF_NOT,                                      // unary
    EF_ADD, EF_SUB, EF_MUL, EF_DIV, EF_MOD,              // binary, operands of `type`
    EF_LT, EF_GT, EF_LE, EF_GE, EF_EQ, EF_NE,            // compare operands of `type`, push int
    EF_AND, EF_OR,                                       // operands compared with 0, push int
    EF_SEL,                                              // c ? a : b (c int; a, b of `type`)
    EF_FABS, EF_SQRT, EF_EXP, EF_LOG, EF_SIN, EF_COS, EF_TAN,
    EF_FLOOR, EF_CEIL, EF_ROUND, EF_TRUNC,               // one argument of `type`
    EF_POW, EF_FMIN, EF_FMAX,                            // two arguments of `type`
    EF_NOPS
};

// One postfix operation: `type` is the type it computes in (and of its result, except for
// comparisons and && || !, which push int).
struct EfIns {
    uint8_t op;
    uint8_t type;
    uint16_t pad_;
    int32_t pad2_;
    double imm;
};

constexpr int kEfMaxIns = 64;    // tape length
constexpr int kEfMaxDepth = 8;   // evaluation stack

// Compiles `text`; false with `err` set for anything outside the supported syntax, a tape
// longer than kEfMaxIns or a stack deeper than kEfMaxDepth.
bool ef_compile(const std::string& text, std::vector<EfIns>* out, std::string* err);

// The host interpreter of a compiled tape (same rules as the device one).
float ef_eval_host(const EfIns* code, int n, float X, float Y);

}  // namespace ppi

#endif
