"""User error functions ERROR(X,Y) (definitions.h:18, CMakeLists.txt:26-33): the engine's compiler
and host interpreter (ppi_ef_evaluate, which the device interpreter follows rule for rule) against
the same expressions compiled by g++ as sequential.cc compiles them, bit for bit; the oracle's
plumbing for them; and what is refused."""
import numpy as np
import pytest

import oracle
import ppiamd
from helpers import Builder, CompiledEF

EXPRS = [
    "fabs((X)-(Y))", "((X)-(Y))*((X)-(Y))", "((X)!=(Y))", "0.5*((X)-(Y))*((X)-(Y))",
    "(((X)-(Y))*((X)-(Y))/(1.0f+fabs(Y)))", "pow((X)-(Y),2)", "fabs(X-Y)/(fabs(Y)+1e-3)",
    "(X)>(Y) ? (X)-(Y) : 2*((Y)-(X))", "sqrt(fabs(X-Y))", "log(1+fabs(X-Y))", "exp(-fabs(X-Y))",
    "(int)(X+0.5f) != (int)Y", "fmin(fabs(X-Y), 1.0)", "fmaxf(X, Y) - fminf(X, Y)", "abs(X-Y)",
    "(X-Y)*(X-Y) + 0.1f*fabsf(X)", "!(X < Y) && (Y >= 0)", "floor(X) == Y", "round(X)-Y", "-(X-Y)",
    "(float)(X-Y)*(double)(X-Y)", "7 % 3 + (X-Y)", "(X != X) ? 1e30 : fabs(X-Y)", "cos(X)*Y",
    "sin(X-Y)*sin(X-Y)", "tan(X)-Y", "(X-Y)/(X+Y)", "X*Y - Y*X + (X-Y)", "sqrtf(fabsf(X)) - truncf(Y)",
    "ceilf(X) <= floorf(Y) || roundf(X) == 2", "(X > 0 ? 1 : -1) * (Y - X)", "(int)X / 3 + (int)Y % 4",
]


def probe_values():
    rng = np.random.default_rng(7)
    special = np.array([0.0, -0.0, 0.5, -0.5, 1.0, -1.0, 1.5, 2.0, -2.5, 3.0, 7.0, 1e-30, 1e-45, 3.0e38, -3.0e38,
                        1e30, np.inf, -np.inf, np.nan, 0.49999997, 2.5, -7.5], dtype=np.float32)
    rnd = (rng.standard_normal(60) * 10.0 ** rng.integers(-3, 6, 60)).astype(np.float32)
    v = np.concatenate([special, rnd])
    X, Y = np.meshgrid(v, v)
    return X.ravel(), Y.ravel()


@pytest.fixture(scope="module")
def refs(tmp_path_factory):
    return CompiledEF(EXPRS, tmp_path_factory.mktemp("ef"))


def same_bits(a, b):
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))


@pytest.mark.parametrize("text", EXPRS)
def test_compiled_expression_matches_gxx(refs, text):
    X, Y = probe_values()
    got = ppiamd.ef_evaluate(text, X, Y)
    want = refs.values(text, X, Y)
    ok = same_bits(got, want)
    assert ok.all(), f"{(~ok).sum()} differ, first at X={X[~ok][0]!r} Y={Y[~ok][0]!r}: {got[~ok][0]!r} vs {want[~ok][0]!r}"


@pytest.mark.parametrize("text", ["expf(X-Y)", "logf(X)", "sinf(X)", "cosf(Y)", "tanf(X)", "powf(X, 2.0f)",
                                  "COST[(int)X][(int)Y]", "Z - Y", "(X - Y", "X - Y)", "fabs(X, Y)",
                                  "X << 2", "1u + X", "+".join(["X"] * 70)])
def test_refused(text):
    with pytest.raises(ppiamd.PpiError):
        ppiamd.ef_evaluate(text, np.zeros(1, np.float32), np.zeros(1, np.float32))


def test_oracle_user_function_matches_builtin(refs):
    """The oracle's EF_EXPR path (a compiled function) gives the built-in error functions' bits for
    their own expressions, under every strategy."""
    b = Builder(4)
    progs = [["+", "x0", 0.5], ["*", "x1", "x2"], ["sin", "x3"], ["/", "x0", "x1"], [2.0], ["exp", "*", 30.0, "x0"]]
    pop = b.population(progs)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 3, 700, 4)
    for text, ef in (("fabs((X)-(Y))", ppiamd.EF_ABS), ("((X)-(Y))*((X)-(Y))", ppiamd.EF_SQUARED),
                     ("((X)!=(Y))", ppiamd.EF_NEQ)):
        for strategy in (ppiamd.SEQ, ppiamd.PP, ppiamd.FP, ppiamd.PDP):
            geo = oracle.reference_geometry(data.shape[0], pop.n)
            a, _ = oracle.interpret(oracle.make_cfg(strategy, ef, False, True, None, *geo), b.table, pop, data, 1e-6)
            c, _ = oracle.interpret(oracle.make_cfg(strategy, ppiamd.EF_EXPR, False, True, None, *geo,
                                                    ef_fn=refs.addr[text]), b.table, pop, data, 1e-6)
            assert np.array_equal(a.view(np.uint32), c.view(np.uint32)), (text, strategy)
