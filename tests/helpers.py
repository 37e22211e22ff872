"""Shared test helpers: hand-written programs, grammars with the full operator set."""
import numpy as np

import ppiamd

ALL_BINARY = ["+", "-", "*", "/", "mean", "max", "min", "mod", "pow", "xor", "=", "!=", "<", "<=", ">", ">=",
              "or", "and"]
ALL_UNARY = ["abs", "sqrt", "pow2", "pow3", "pow4", "pow5", "neg", "round", "ceil", "floor", "exp", "exp10", "exp2",
             "log", "log10", "log2", "sin", "cos", "tan", "step", "sign", "logistic", "gamma", "gauss", "not"]
CONFIG_BINARY = ppiamd.BINARY
CONFIG_UNARY = ppiamd.UNARY


class Builder:
    """Builds phenotype arrays from prefix token lists, e.g. ['-', 'x0', 1.5]."""

    def __init__(self, n_attributes=4, extra_terminals=()):
        bnf_terms = " | ".join([f"attr{k}" for k in range(n_attributes)] + ["const"] + list(extra_terminals))
        bnf = ("S = <e>\nP = <e> ::= <b> <e> <e> | <u> <e> | if_then_else <e> <e> <e> | <t>\n"
               f"    <b> ::= {' | '.join(ALL_BINARY)}\n    <u> ::= {' | '.join(ALL_UNARY)}\n    <t> ::= {bnf_terms}\n")
        g = ppiamd._grammar.parse_bnf(bnf)
        self.values = g.values()
        self.table = g.engine_table()
        self.t_attr = self.values["T_ATTRIBUTE"]
        self.t_const = self.values["T_CONST"]

    def sym(self, tok):
        if tok == "ite":
            return self.values["T_IF_THEN_ELSE"]
        return self.values[ppiamd._grammar._terminal_name(tok)]

    def encode(self, tokens):
        ph, ep = [], []
        for t in tokens:
            if isinstance(t, (float, np.floating)):
                ph.append(self.t_const); ep.append(float(t))
            elif isinstance(t, str) and t.startswith("x") and t[1:].isdigit():
                ph.append(self.t_attr); ep.append(float(int(t[1:])))
            elif isinstance(t, tuple) and t[0] == "sym":
                ph.append(int(t[1])); ep.append(float(t[2]) if len(t) > 2 else 0.0)
            else:
                ph.append(self.sym(t)); ep.append(0.0)
        return ph, ep

    def population(self, programs, stride=None):
        enc = [self.encode(p) for p in programs]
        stride = stride or max(1, max(len(p) for p, _ in enc))
        n = len(enc)
        ph = np.zeros(n * stride, dtype=np.int32)
        ep = np.zeros(n * stride, dtype=np.float32)
        sz = np.zeros(n, dtype=np.int32)
        for i, (p, e) in enumerate(enc):
            ph[i * stride:i * stride + len(p)] = p
            ep[i * stride:i * stride + len(e)] = e
            sz[i] = len(p)
        return ppiamd.Population(ph, ep, sz, stride)


def full_grammar(n_attributes):
    return ppiamd.Grammar(n_attributes, binary=ALL_BINARY, unary=ALL_UNARY)


def rel_close(a, b, rtol):
    """Elementwise |a-b| <= rtol*max(|a|,|b|) with exact equality for non-finite / huge sentinels."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(a), np.abs(b))
    ok = (a == b) | (np.abs(a - b) <= rtol * scale)
    return ok


def sums_agree(vec, ovec, exact_sum, bad, nlin, alpha, size, rtol=1e-5):
    """Per-program agreement bar for float error sums whose reduction order differs.

    A value passes if it is within rtol of the oracle's, or if it is at least as
    close to the exact sum of the same per-row float errors (exact_sum / nlin +
    alpha * size, computed in double) as the oracle's own sequential float sum is.
    Programs with a non-finite row error or size 0 (FLT_MAX-based sentinels) and
    sums that overflow must match exactly.
    """
    vec = np.asarray(vec, np.float64)
    ovec = np.asarray(ovec, np.float64)
    size = np.asarray(size, np.float64)
    ex = np.asarray(exact_sum, np.float64) / nlin + np.float64(np.float32(alpha)) * size
    close = rel_close(vec, ovec, rtol)
    err_g = np.abs(vec - ex)
    err_o = np.abs(ovec - ex)
    better = err_g <= err_o + 1e-6 * np.abs(ex)
    sentinel = (np.asarray(bad) != 0) | (size == 0) | (ovec >= 3.4e38) | (vec >= 3.4e38)
    return np.where(sentinel, vec == ovec, close | better)


# The headers sequential.cc includes (sequential.cc:16-22): a user ERROR(X,Y) compiled after them
# sees the overloads the reference's sequential path sees.
_EF_PRELUDE = ("#include <stdio.h>\n#include <stdlib.h>\n#include <iostream>\n#include <cmath>\n#include <string>\n"
               "#include <limits>\n#include <queue>\n")


class CompiledEF:
    """User error functions compiled by g++ as the reference compiles them into sequential.cc
    (`float error = ERROR(X, Y);`, no contraction): `fn[text]` is a ctypes function
    float(float, float) and `addr[text]` its address (for the oracle's ef_fn)."""

    def __init__(self, texts, workdir):
        import ctypes as C
        import os
        import subprocess
        src = os.path.join(str(workdir), "ef_refs.cc")
        lib = os.path.join(str(workdir), "libef_refs.so")
        with open(src, "w") as f:
            f.write(_EF_PRELUDE)
            for k, t in enumerate(texts):
                f.write(f'extern "C" float ef_ref_{k}(float X, float Y) {{ float error = {t}; return error; }}\n')
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", lib, src],
                       check=True)
        self._lib = C.CDLL(lib)
        self.fn, self.addr = {}, {}
        for k, t in enumerate(texts):
            f = getattr(self._lib, f"ef_ref_{k}")
            f.restype = C.c_float
            f.argtypes = [C.c_float, C.c_float]
            self.fn[t] = f
            self.addr[t] = C.cast(f, C.c_void_p).value

    def values(self, text, X, Y):
        f = self.fn[text]
        return np.array([f(float(x), float(y)) for x, y in zip(X, Y)], dtype=np.float32)
