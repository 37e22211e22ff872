"""GPU parity of the ordered form: sums bit-identical to the sequential path.

The ordered form evaluates row tiles like the FP form, stores every row's error,
and sums each program's errors in row order with the sequential loop's rules
(sequential.cc:169-179: float accumulation, stop at the first non-finite error,
REDUCEMAX max(EF*nlin)).  Bar: every result bit-identical to the oracle's, for
SEQ and PP semantics, in one pass and in several passes over the rows.
"""
import numpy as np
import pytest

import oracle
import ppiamd
from helpers import Builder, full_grammar

pytestmark = pytest.mark.gpu

ALPHA = np.float32(1e-6)


@pytest.fixture(scope="module")
def eng():
    if ppiamd.device_count() == 0:
        pytest.fail("no HIP device visible")
    e = ppiamd.Engine(0)
    yield e
    e.close()


def ordered_vs_oracle(eng, table, pop, data, error_fn, strategy=ppiamd.SEQ, reducemax=False, protected=True,
                      cost=None, pass_rows=0):
    eng.configure(table, strategy=strategy, error_fn=error_fn, reducemax=reducemax, protected=protected, cost=cost,
                  form=ppiamd.FORM_ORDERED, ordered_pass_rows=pass_rows, compute_units=256)
    eng.set_dataset(data)
    vec, idx = eng.evaluate(pop, alpha=ALPHA, best_size=1)
    assert eng.stats().form == ppiamd.FORM_ORDERED
    dpl, ppl, bpl = oracle.reference_geometry(data.shape[0], pop.n)
    cfg = oracle.make_cfg(strategy, error_fn, reducemax, protected, cost, dpl, ppl, bpl)
    ovec, oidx = oracle.interpret(cfg, table, pop, data, alpha=ALPHA, best_size=1)
    return vec, idx, ovec, oidx


def assert_bits(vec, ovec, what):
    same = vec.view(np.uint32) == ovec.view(np.uint32)
    bad = np.nonzero(~same)[0]
    assert bad.size == 0, f"{what}: {bad.size} differ, e.g. programs {bad[:5]}: {vec[bad[:5]]} vs {ovec[bad[:5]]}"


@pytest.mark.parametrize("pass_rows", [0, 3000])
@pytest.mark.parametrize("strategy", [ppiamd.SEQ, ppiamd.PP])
def test_regression_bit_exact(eng, strategy, pass_rows):
    # 30,011 rows: the last row tile is partial; 3,000 rows per pass gives 11 passes.
    w = ppiamd.workload("regression", n_programs=2000, nlin=30011, seed=5)
    print(f"case ordered regression strategy={strategy} pass_rows={pass_rows}", flush=True)
    vec, idx, ovec, oidx = ordered_vs_oracle(eng, w.grammar.table, w.pop, w.data, w.error_fn, strategy=strategy,
                                             pass_rows=pass_rows)
    assert_bits(vec, ovec, "regression")
    assert np.array_equal(idx, oidx)


@pytest.mark.parametrize("reducemax", [False, True])
def test_classification_bit_exact(eng, reducemax):
    w = ppiamd.workload("classification", n_programs=2000, nlin=20000, seed=8)
    print(f"case ordered classification rm={reducemax}", flush=True)
    vec, idx, ovec, oidx = ordered_vs_oracle(eng, w.grammar.table, w.pop, w.data, w.error_fn, reducemax=reducemax,
                                             cost=w.cost, pass_rows=4096)
    assert_bits(vec, ovec, "classification")
    assert np.array_equal(idx, oidx)


@pytest.mark.parametrize("prot", [True, False])
def test_full_operator_set_bit_exact(eng, prot):
    """Every operator, with rows whose error overflows (the sum stops there)."""
    gr = full_grammar(6)
    pop = ppiamd.grow(gr, 1200, 11, 7, 96, p_leaf=0.3)
    data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 12, 3000, 6) * np.float32(3.0)
    print(f"case ordered full-ops protected={prot}", flush=True)
    for ef in (ppiamd.EF_ABS, ppiamd.EF_SQUARED):
        vec, idx, ovec, oidx = ordered_vs_oracle(eng, gr.table, pop, data, ef, protected=prot, pass_rows=1000)
        assert_bits(vec, ovec, f"full operator set, error function {ef}")
    assert (ovec >= np.float32(3.4e38) / data.shape[0]).any(), "the population must exercise non-finite rows"


def test_edge_programs_bit_exact(eng):
    b = Builder(4)
    progs = [["-", "x0", "x1"], ["/", "x0", 0.0], ["log", 2.0], ["ite", ">", "x0", 0.0, "x1", "x2"],
             ["x0", "x1", "x2"], ["+", "x0"], [("sym", 4242)], [("sym", b.t_attr, 9.0)],
             ["exp", "exp", "exp", "x0"], ["mod", "x0", 0.0], ["gamma", "abs", "x1"]]
    pop = b.population(progs)
    pop.size[4] = 0  # an empty program
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 5, 777, 4)
    for strategy in (ppiamd.SEQ, ppiamd.PP):
        print(f"case ordered edge strategy={strategy}", flush=True)
        vec, idx, ovec, oidx = ordered_vs_oracle(eng, b.table, pop, data, ppiamd.EF_ABS, strategy=strategy)
        assert_bits(vec, ovec, "edge programs")
        assert np.array_equal(idx, oidx)


def test_rejects_tree_strategies(eng):
    b = Builder(2)
    for strategy in (ppiamd.FP, ppiamd.PDP):
        with pytest.raises(ppiamd.PpiError):
            eng.configure(b.table, strategy=strategy, form=ppiamd.FORM_ORDERED)


@pytest.mark.parametrize("rows_per_lane", [1, 2, 4, 8])
def test_planted_rows_bit_exact(eng, rows_per_lane):
    """Program k returns column k; the target is 0, so the row errors are the column's values.
    Each column plants a non-finite value (or a run of huge ones) at a chosen row: the first row,
    both ends of a 64-row batch, a row tile boundary, mid-batch, the partial last tile, a later
    pass; sums that overflow with finite errors only; -0 errors.  Sum and REDUCEMAX, one pass and
    several: bit-identical to the oracle's sequential loop."""
    nlin = 5000 + 37  # the last row tile is partial at every rows per lane
    rng = np.random.default_rng(rows_per_lane)
    plants = [(0, np.inf), (63, np.nan), (64, -np.inf), (127, np.nan), (255, np.inf), (256, np.nan),
              (511, np.inf), (100, np.nan), (nlin - 1, np.inf), (nlin - 20, np.nan), (4000, np.inf),
              (None, None), (None, "huge"), (None, "negzero")]
    cols = []
    for row, v in plants:
        c = rng.uniform(0.0, 2.0, nlin).astype(np.float32)
        if v == "huge":
            c[1000:1200] = np.float32(3.0e38)  # finite errors whose row-order sum overflows
        elif v == "negzero":
            c[:] = np.float32(-0.0)
        elif row is not None:
            c[row] = np.float32(v)
        cols.append(c)
    data = np.stack(cols + [np.zeros(nlin, np.float32)], axis=1)
    b = Builder(len(cols))
    pop = b.population([[f"x{k}"] for k in range(len(cols))] + [["+", f"x{k}", 1.0] for k in range(len(cols))])
    for reducemax in (False, True):
        for pass_rows in (0, 1024):
            eng.configure(b.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_ABS, reducemax=reducemax,
                          form=ppiamd.FORM_ORDERED, ordered_pass_rows=pass_rows, rows_per_lane=rows_per_lane,
                          compute_units=256)
            eng.set_dataset(data)
            vec, idx = eng.evaluate(pop, alpha=ALPHA, best_size=1)
            dpl, ppl, bpl = oracle.reference_geometry(nlin, pop.n)
            cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS, reducemax, True, None, dpl, ppl, bpl)
            ovec, oidx = oracle.interpret(cfg, b.table, pop, data, alpha=ALPHA, best_size=1)
            print(f"case planted R={rows_per_lane} rm={reducemax} pass_rows={pass_rows}", flush=True)
            assert_bits(vec, ovec, f"planted rows R={rows_per_lane} rm={reducemax} pass_rows={pass_rows}")
            assert np.array_equal(idx, oidx)
