// ef.cc — compiler and host interpreter of user error functions (ef.h).
#include "ef.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

namespace ppi {
namespace {

struct Node {
    uint8_t op = EF_X;       // EF_* (EF_IMM for literals, EF_CVT for casts)
    uint8_t type = EF_FLOAT;  // type of the result
    uint8_t ctype = EF_FLOAT; // type the operation computes in (its operands' type)
    double imm = 0.0;
    std::vector<std::unique_ptr<Node>> kid;
};
typedef std::unique_ptr<Node> NodeP;

struct Func {
    const char* name;
    uint8_t op;
    int nargs;
    int type;  // EF_FLOAT, EF_DOUBLE, or -1: the argument's own type (abs)
};
// The float routines of the C library that are exact (IEEE operations and roundings) are
// accepted; its float transcendentals (expf, logf, sinf, cosf, tanf, powf) are not correctly
// rounded, so their bits cannot be reproduced from their definition, and are refused.
const Func kFuncs[] = {
    {"fabs", EF_FABS, 1, EF_DOUBLE},   {"fabsf", EF_FABS, 1, EF_FLOAT},   {"abs", EF_FABS, 1, -1},
    {"sqrt", EF_SQRT, 1, EF_DOUBLE},   {"sqrtf", EF_SQRT, 1, EF_FLOAT},   {"exp", EF_EXP, 1, EF_DOUBLE},
    {"log", EF_LOG, 1, EF_DOUBLE},     {"sin", EF_SIN, 1, EF_DOUBLE},     {"cos", EF_COS, 1, EF_DOUBLE},
    {"tan", EF_TAN, 1, EF_DOUBLE},     {"floor", EF_FLOOR, 1, EF_DOUBLE}, {"floorf", EF_FLOOR, 1, EF_FLOAT},
    {"ceil", EF_CEIL, 1, EF_DOUBLE},   {"ceilf", EF_CEIL, 1, EF_FLOAT},   {"round", EF_ROUND, 1, EF_DOUBLE},
    {"roundf", EF_ROUND, 1, EF_FLOAT}, {"trunc", EF_TRUNC, 1, EF_DOUBLE}, {"truncf", EF_TRUNC, 1, EF_FLOAT},
    {"pow", EF_POW, 2, EF_DOUBLE},     {"fmin", EF_FMIN, 2, EF_DOUBLE},   {"fminf", EF_FMIN, 2, EF_FLOAT},
    {"fmax", EF_FMAX, 2, EF_DOUBLE},   {"fmaxf", EF_FMAX, 2, EF_FLOAT},
};
const char* const kRefused[] = {"expf", "logf", "sinf", "cosf", "tanf", "powf"};

NodeP leaf(uint8_t op, uint8_t type, double imm = 0.0) {
    NodeP n(new Node);
    n->op = op;
    n->type = type;
    n->ctype = type;
    n->imm = imm;
    return n;
}

NodeP convert(NodeP a, uint8_t type) {
    if (a->type == type) return a;
    NodeP n = leaf(EF_CVT, type);
    n->kid.push_back(std::move(a));
    return n;
}

struct Parser {
    const std::string& s;
    size_t i = 0;
    std::string err;
    explicit Parser(const std::string& t) : s(t) {}

    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool eat(const char* tok) {
        ws();
        const size_t n = std::strlen(tok);
        if (s.compare(i, n, tok) != 0) return false;
        // "<" must not take the first character of "<=", and so on
        if (n == 1 && i + 1 < s.size() && std::strchr("<>=!&|", tok[0]) && s[i + 1] == '=' && tok[0] != '=') return false;
        if (n == 1 && (tok[0] == '&' || tok[0] == '|') && i + 1 < s.size() && s[i + 1] == tok[0]) return false;
        i += n;
        return true;
    }
    bool fail(const std::string& m) {
        if (err.empty()) err = m + " at offset " + std::to_string(i) + " of \"" + s + "\"";
        return false;
    }
    std::string ident() {
        ws();
        size_t j = i;
        while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) ++j;
        return s.substr(i, j - i);
    }

    NodeP binary(uint8_t op, NodeP a, NodeP b) {
        const uint8_t t = std::max(a->type, b->type);  // the usual arithmetic conversions
        NodeP n = leaf(op, t);
        if (op >= EF_LT && op <= EF_NE) n->type = EF_INT;  // comparisons give int
        n->kid.push_back(convert(std::move(a), t));
        n->kid.push_back(convert(std::move(b), t));
        return n;
    }
    NodeP to_double(NodeP a) { return convert(std::move(a), EF_DOUBLE); }

    NodeP primary() {
        ws();
        if (i >= s.size()) { fail("unexpected end"); return nullptr; }
        const char c = s[i];
        if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < s.size() && std::isdigit((unsigned char)s[i + 1]))) {
            size_t j = i;
            bool is_int = true;
            if (s.compare(i, 2, "0x") == 0 || s.compare(i, 2, "0X") == 0) {
                j += 2;
                while (j < s.size() && std::isxdigit((unsigned char)s[j])) ++j;
            } else {
                while (j < s.size() && (std::isdigit((unsigned char)s[j]) || s[j] == '.' || s[j] == 'e' || s[j] == 'E' ||
                                        ((s[j] == '+' || s[j] == '-') && (s[j - 1] == 'e' || s[j - 1] == 'E')))) {
                    if (s[j] == '.' || s[j] == 'e' || s[j] == 'E') is_int = false;
                    ++j;
                }
            }
            const std::string lit = s.substr(i, j - i);
            NodeP n;
            if (j < s.size() && (s[j] == 'f' || s[j] == 'F') && !is_int) {
                n = leaf(EF_IMM, EF_FLOAT, (double)std::strtof(lit.c_str(), nullptr));
                ++j;
            } else if (is_int) {
                const long long v = std::strtoll(lit.c_str(), nullptr, 0);
                if (v > 2147483647LL) { fail("integer literal beyond int"); return nullptr; }
                n = leaf(EF_IMM, EF_INT, (double)v);
                while (j < s.size() && (s[j] == 'u' || s[j] == 'U' || s[j] == 'l' || s[j] == 'L')) {
                    fail("integer suffixes are not supported");
                    return nullptr;
                }
            } else {
                n = leaf(EF_IMM, EF_DOUBLE, std::strtod(lit.c_str(), nullptr));
            }
            i = j;
            return n;
        }
        if (eat("(")) {
            NodeP e = expr();
            if (!e) return nullptr;
            if (!eat(")")) { fail("expected )"); return nullptr; }
            return e;
        }
        const std::string id = ident();
        if (id.empty()) { fail("unexpected character"); return nullptr; }
        i += id.size();
        if (id == "X") return leaf(EF_X, EF_FLOAT);
        if (id == "Y") return leaf(EF_Y, EF_FLOAT);
        for (const Func& f : kFuncs) {
            if (id != f.name) continue;
            if (!eat("(")) { fail("expected ( after " + id); return nullptr; }
            std::vector<NodeP> args;
            for (int a = 0; a < f.nargs; ++a) {
                if (a > 0 && !eat(",")) { fail("expected , in " + id); return nullptr; }
                NodeP e = cond();
                if (!e) return nullptr;
                args.push_back(std::move(e));
            }
            if (!eat(")")) { fail("expected ) after the arguments of " + id); return nullptr; }
            uint8_t t = f.type < 0 ? args[0]->type : (uint8_t)f.type;
            NodeP n = leaf(f.op, t);
            for (auto& e : args) n->kid.push_back(convert(std::move(e), t));
            return n;
        }
        for (const char* r : kRefused)
            if (id == r) {
                fail(id + ": the C library's float transcendentals are not correctly rounded and are not "
                          "reproduced; use the double routine (" + id.substr(0, id.size() - 1) + ")");
                return nullptr;
            }
        fail("unknown name " + id);
        return nullptr;
    }

    NodeP unary() {
        ws();
        if (eat("-")) {
            NodeP a = unary();
            if (!a) return nullptr;
            NodeP n = leaf(EF_NEG, a->type);
            n->kid.push_back(std::move(a));
            return n;
        }
        if (eat("+")) return unary();
        if (eat("!")) {
            NodeP a = unary();
            if (!a) return nullptr;
            NodeP n = leaf(EF_NOT, EF_INT);
            n->ctype = EF_DOUBLE;
            n->kid.push_back(to_double(std::move(a)));
            return n;
        }
        // cast
        const size_t save = i;
        if (eat("(")) {
            const std::string id = ident();
            uint8_t t = 255;
            if (id == "int") t = EF_INT;
            else if (id == "float") t = EF_FLOAT;
            else if (id == "double") t = EF_DOUBLE;
            if (t != 255) {
                i += id.size();
                if (eat(")")) {
                    NodeP a = unary();
                    if (!a) return nullptr;
                    NodeP n = leaf(EF_CVT, t);
                    n->kid.push_back(std::move(a));
                    return n;
                }
            }
            i = save;
        }
        return primary();
    }
    NodeP mul() {
        NodeP a = unary();
        while (a) {
            uint8_t op;
            if (eat("*")) op = EF_MUL;
            else if (eat("/")) op = EF_DIV;
            else if (eat("%")) op = EF_MOD;
            else break;
            NodeP b = unary();
            if (!b) return nullptr;
            if (op == EF_MOD && (a->type != EF_INT || b->type != EF_INT)) { fail("% needs int operands"); return nullptr; }
            a = binary(op, std::move(a), std::move(b));
        }
        return a;
    }
    NodeP add() {
        NodeP a = mul();
        while (a) {
            uint8_t op;
            if (eat("+")) op = EF_ADD;
            else if (eat("-")) op = EF_SUB;
            else break;
            NodeP b = mul();
            if (!b) return nullptr;
            a = binary(op, std::move(a), std::move(b));
        }
        return a;
    }
    NodeP rel() {
        NodeP a = add();
        while (a) {
            uint8_t op;
            if (eat("<=")) op = EF_LE;
            else if (eat(">=")) op = EF_GE;
            else if (eat("<")) op = EF_LT;
            else if (eat(">")) op = EF_GT;
            else break;
            NodeP b = add();
            if (!b) return nullptr;
            NodeP n = binary(op, std::move(a), std::move(b));
            a = std::move(n);
        }
        return a;
    }
    NodeP eq() {
        NodeP a = rel();
        while (a) {
            uint8_t op;
            if (eat("==")) op = EF_EQ;
            else if (eat("!=")) op = EF_NE;
            else break;
            NodeP b = rel();
            if (!b) return nullptr;
            a = binary(op, std::move(a), std::move(b));
        }
        return a;
    }
    NodeP land() {
        NodeP a = eq();
        while (a && eat("&&")) {
            NodeP b = eq();
            if (!b) return nullptr;
            NodeP n = leaf(EF_AND, EF_INT);
            n->ctype = EF_DOUBLE;
            n->kid.push_back(to_double(std::move(a)));
            n->kid.push_back(to_double(std::move(b)));
            a = std::move(n);
        }
        return a;
    }
    NodeP lor() {
        NodeP a = land();
        while (a && eat("||")) {
            NodeP b = land();
            if (!b) return nullptr;
            NodeP n = leaf(EF_OR, EF_INT);
            n->ctype = EF_DOUBLE;
            n->kid.push_back(to_double(std::move(a)));
            n->kid.push_back(to_double(std::move(b)));
            a = std::move(n);
        }
        return a;
    }
    NodeP cond() {
        NodeP c = lor();
        if (!c || !eat("?")) return c;
        NodeP a = expr();
        if (!a) return nullptr;
        if (!eat(":")) { fail("expected : in ?:"); return nullptr; }
        NodeP b = cond();
        if (!b) return nullptr;
        const uint8_t t = std::max(a->type, b->type);
        NodeP n = leaf(EF_SEL, t);
        n->kid.push_back(to_double(std::move(c)));
        n->kid.push_back(convert(std::move(a), t));
        n->kid.push_back(convert(std::move(b), t));
        return n;
    }
    NodeP expr() { return cond(); }
};

// Postfix: operands first, then the operation tagged with the type it computes in.
void emit(const Node& n, std::vector<EfIns>* out, int depth, int* max_depth) {
    int d = depth;
    for (const auto& k : n.kid) {
        emit(*k, out, d, max_depth);
        ++d;
    }
    EfIns e{};
    e.op = n.op;
    e.type = n.ctype;
    e.imm = n.imm;
    out->push_back(e);
    *max_depth = std::max(*max_depth, std::max(d, depth + 1));
}

// --- host evaluation -------------------------------------------------------
int32_t to_int(double v) {
    // C truncation; out of range and NaN give INT_MIN (the x86 conversion's "indefinite" value)
    if (!(v > -2147483649.0 && v < 2147483648.0)) return INT32_MIN;
    return (int32_t)v;
}
double round_to(double v, uint8_t type) {
    if (type == EF_FLOAT) return (double)(float)v;
    if (type == EF_INT) return (double)to_int(v);
    return v;
}
int32_t wrap(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }

}  // namespace

bool ef_compile(const std::string& text, std::vector<EfIns>* out, std::string* err) {
    out->clear();
    Parser p(text);
    NodeP root = p.expr();
    p.ws();
    if (root && p.i != text.size()) p.fail("unexpected text");
    if (!root || !p.err.empty()) {
        if (err) *err = p.err.empty() ? "cannot parse the error function" : p.err;
        return false;
    }
    root = convert(std::move(root), EF_FLOAT);  // float error = ERROR(X, Y)
    int depth = 0;
    emit(*root, out, 0, &depth);
    if ((int)out->size() > kEfMaxIns || depth > kEfMaxDepth) {
        if (err) *err = "error function too long (at most " + std::to_string(kEfMaxIns) + " operations, stack depth " +
                        std::to_string(kEfMaxDepth) + ")";
        out->clear();
        return false;
    }
    return true;
}

float ef_eval_host(const EfIns* code, int n, float X, float Y) {
    double st[kEfMaxDepth];
    int sp = 0;
    for (int k = 0; k < n; ++k) {
        const EfIns& e = code[k];
        const uint8_t t = e.type;
        switch (e.op) {
        case EF_X: st[sp++] = X; break;
        case EF_Y: st[sp++] = Y; break;
        case EF_IMM: st[sp++] = e.imm; break;
        case EF_CVT: st[sp - 1] = round_to(st[sp - 1], t); break;
        case EF_NEG: st[sp - 1] = t == EF_INT ? (double)wrap(-(int64_t)st[sp - 1]) : -st[sp - 1]; break;
        case EF_NOT: st[sp - 1] = st[sp - 1] == 0.0 ? 1.0 : 0.0; break;
        case EF_FABS: st[sp - 1] = t == EF_INT ? (double)wrap(std::llabs((int64_t)st[sp - 1])) : std::fabs(st[sp - 1]); break;
        case EF_SQRT: st[sp - 1] = round_to(std::sqrt(st[sp - 1]), t); break;
        case EF_EXP: st[sp - 1] = round_to(std::exp(st[sp - 1]), t); break;
        case EF_LOG: st[sp - 1] = round_to(std::log(st[sp - 1]), t); break;
        case EF_SIN: st[sp - 1] = round_to(std::sin(st[sp - 1]), t); break;
        case EF_COS: st[sp - 1] = round_to(std::cos(st[sp - 1]), t); break;
        case EF_TAN: st[sp - 1] = round_to(std::tan(st[sp - 1]), t); break;
        case EF_FLOOR: st[sp - 1] = std::floor(st[sp - 1]); break;
        case EF_CEIL: st[sp - 1] = std::ceil(st[sp - 1]); break;
        case EF_ROUND: st[sp - 1] = std::round(st[sp - 1]); break;
        case EF_TRUNC: st[sp - 1] = std::trunc(st[sp - 1]); break;
        case EF_SEL: {
            const double b = st[--sp], a = st[--sp], c = st[--sp];
            st[sp++] = c != 0.0 ? a : b;
            break;
        }
        default: {
            const double b = st[--sp], a = st[--sp];
            double r = 0.0;
            switch (e.op) {
            case EF_ADD: r = t == EF_INT ? (double)wrap((int64_t)a + (int64_t)b) : round_to(a + b, t); break;
            case EF_SUB: r = t == EF_INT ? (double)wrap((int64_t)a - (int64_t)b) : round_to(a - b, t); break;
            case EF_MUL: r = t == EF_INT ? (double)wrap((int64_t)a * (int64_t)b) : round_to(a * b, t); break;
            case EF_DIV:
                if (t == EF_INT) r = b == 0.0 ? 0.0 : (double)wrap((int64_t)a / (int64_t)b);
                else r = round_to(a / b, t);
                break;
            case EF_MOD: r = b == 0.0 ? 0.0 : (double)wrap((int64_t)a % (int64_t)b); break;
            case EF_LT: r = a < b; break;
            case EF_GT: r = a > b; break;
            case EF_LE: r = a <= b; break;
            case EF_GE: r = a >= b; break;
            case EF_EQ: r = a == b; break;
            case EF_NE: r = a != b; break;
            case EF_AND: r = (a != 0.0) && (b != 0.0); break;
            case EF_OR: r = (a != 0.0) || (b != 0.0); break;
            case EF_POW: r = round_to(std::pow(a, b), t); break;
            case EF_FMIN: r = std::fmin(a, b); break;
            case EF_FMAX: r = std::fmax(a, b); break;
            default: r = 0.0; break;
            }
            st[sp++] = r;
            break;
        }
        }
    }
    return (float)st[sp - 1];
}

}  // namespace ppi
