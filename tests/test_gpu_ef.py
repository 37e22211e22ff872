"""GPU: user error functions (PPI_EF_EXPR) through the engine against the oracle, which calls the
same expressions compiled by g++ as sequential.cc compiles them.  Every program takes the
reference's own partition and row order for the strategy, so arithmetic expressions give the
oracle's bits; expressions through the double transcendentals use the device library's double
routines (within one double ulp of glibc's), whose float-rounded results can differ in the last
bit on rare rows: those sums are held to the 1e-5 relative bar."""
import numpy as np
import pytest

import oracle
import ppiamd
from helpers import Builder, CompiledEF, rel_close

pytestmark = pytest.mark.gpu
ALPHA = np.float32(1e-6)
EXACT = ["(((X)-(Y))*((X)-(Y))/(1.0f+fabs(Y)))", "(X)>(Y) ? (X)-(Y) : 2*((Y)-(X))", "fmin(fabs(X-Y), 1.0)",
         "(int)(X+0.5f) != (int)Y", "0.5*((X)-(Y))*((X)-(Y)) + 0.25f*fabsf(X)"]
TRANSCENDENTAL = ["log(1+fabs(X-Y))", "pow((X)-(Y),2)", "1.0 - exp(-fabs(X-Y))"]


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    gr = ppiamd.Grammar(4)
    pop = ppiamd.grow(gr, 700, 41, 6, 64)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 42, 2100, 4)
    return gr, pop, data, CompiledEF(EXACT + TRANSCENDENTAL, tmp_path_factory.mktemp("efgpu"))


@pytest.mark.parametrize("strategy", [ppiamd.SEQ, ppiamd.PP, ppiamd.FP, ppiamd.PDP])
@pytest.mark.parametrize("text", EXACT + TRANSCENDENTAL)
def test_user_error_function_matches_oracle(case, strategy, text):
    gr, pop, data, ref = case
    print(f"user EF {text!r} strategy {strategy}", flush=True)
    e = ppiamd.Engine(0)
    e.configure(gr.table, strategy=strategy, protected=True, ef_expr=text)
    e.set_dataset(data)
    vec, idx = e.evaluate(pop, alpha=ALPHA, best_size=3)
    e.close()
    geo = oracle.reference_geometry(data.shape[0], pop.n)
    cfg = oracle.make_cfg(strategy, ppiamd.EF_EXPR, False, True, None, *geo, ef_fn=ref.addr[text])
    ovec, oidx = oracle.interpret(cfg, gr.table, pop, data, alpha=ALPHA, best_size=3)
    if text in EXACT:
        assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32)), \
            f"{int((vec.view(np.uint32) != ovec.view(np.uint32)).sum())} differ"
        assert list(idx) == list(oidx)
    else:
        same = (vec.view(np.uint32) == ovec.view(np.uint32))
        assert (same | rel_close(vec, ovec, 1e-5)).all()
        assert same.mean() > 0.95


@pytest.mark.parametrize("strategy", [ppiamd.SEQ, ppiamd.PP, ppiamd.FP, ppiamd.PDP])
def test_user_error_function_long_dataset(case, strategy):
    """12,000 rows: the programs go through the recheck's deep list (16 at a time, row errors computed
    across the device) and the rest through the recheck kernel; bit for bit either way."""
    gr, _, _, ref = case
    pop = ppiamd.grow(gr, 40, 44, 6, 64)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 45, 12000, 4)
    text = EXACT[0]
    print(f"user EF long dataset strategy {strategy}", flush=True)
    e = ppiamd.Engine(0)
    e.configure(gr.table, strategy=strategy, protected=True, ef_expr=text)
    e.set_dataset(data)
    vec, idx = e.evaluate(pop, alpha=ALPHA, best_size=3)
    e.close()
    geo = oracle.reference_geometry(data.shape[0], pop.n)
    cfg = oracle.make_cfg(strategy, ppiamd.EF_EXPR, False, True, None, *geo, ef_fn=ref.addr[text])
    ovec, oidx = oracle.interpret(cfg, gr.table, pop, data, alpha=ALPHA, best_size=3)
    assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32)), \
        f"{int((vec.view(np.uint32) != ovec.view(np.uint32)).sum())} differ"
    assert list(idx) == list(oidx)


def test_user_error_function_reducemax_and_sentinels(case):
    """REDUCEMAX and programs whose error overflows (sentinels) through a user expression, bit for bit."""
    b = Builder(4)
    progs = [["exp", "*", 90.0, "x0"], ["+", "x0", "x1"], [], ["/", "x0", 0.0], ["*", 1e20, "*", 1e20, "x2"]]
    pop = b.population(progs)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 43, 1500, 4)
    text = "(((X)-(Y))*((X)-(Y))/(1.0f+fabs(Y)))"
    ref = case[3]
    for strategy in (ppiamd.SEQ, ppiamd.FP, ppiamd.PDP):
        for rmax in (False, True):
            e = ppiamd.Engine(0)
            e.configure(b.table, strategy=strategy, protected=True, reducemax=rmax, ef_expr=text)
            e.set_dataset(data)
            vec, _ = e.evaluate(pop, alpha=ALPHA)
            e.close()
            geo = oracle.reference_geometry(data.shape[0], pop.n)
            cfg = oracle.make_cfg(strategy, ppiamd.EF_EXPR, rmax, True, None, *geo, ef_fn=ref.addr[text])
            ovec, _ = oracle.interpret(cfg, b.table, pop, data, alpha=ALPHA)
            assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32)), (strategy, rmax, vec, ovec)


def test_user_error_function_refused_where_unsupported(case):
    gr, pop, data, _ = case
    with pytest.raises(ppiamd.PpiError):
        e = ppiamd.Engine(0)
        try:
            e.configure(gr.table, strategy=ppiamd.SEQ, ef_expr="expf(X) - Y")
        finally:
            e.close()
    b = ppiamd.Batch(0)
    with pytest.raises(ppiamd.PpiError):
        b.configure(gr.table, strategy=ppiamd.SEQ, ef_expr="fabs(X-Y)")
    b.close()
