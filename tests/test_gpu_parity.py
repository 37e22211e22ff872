"""GPU parity: the HIP engine against the CPU oracle on the same seeded inputs.

Bars (stated per check): bit-exact for COST/classification counts, for every
non-finite/empty-program sentinel and for the PP form (which keeps the
reference's row order); relative 1e-5 for float sums whose reduction order
differs from the reference's.
"""
import numpy as np
import pytest

import oracle
import ppiamd
from helpers import Builder, full_grammar, rel_close, sums_agree

pytestmark = pytest.mark.gpu

ALPHA = np.float32(1e-6)
RTOL = 1e-5


@pytest.fixture(scope="module")
def eng():
    if ppiamd.device_count() == 0:
        pytest.fail("no HIP device visible")
    e = ppiamd.Engine(0)
    yield e
    e.close()


def run_both(eng, w, strategy, form, alpha=ALPHA, best_size=1, reducemax=None, native=False, rows_per_lane=0,
             programs_per_tile=0, data=None, pop=None, table=None):
    data = w.data if data is None else data
    pop = w.pop if pop is None else pop
    table = w.grammar.table if table is None else table
    rm = w.reducemax if reducemax is None else reducemax
    eng.configure(table, strategy=strategy, error_fn=w.error_fn, reducemax=rm, protected=w.protected,
                  cost=w.cost, form=form, native=native, rows_per_lane=rows_per_lane,
                  programs_per_tile=programs_per_tile, compute_units=256)
    eng.set_dataset(data)
    vec, idx = eng.evaluate(pop, alpha=alpha, best_size=best_size)
    dpl, ppl, bpl = oracle.reference_geometry(data.shape[0], pop.n)
    cfg = oracle.make_cfg(strategy, w.error_fn, rm, w.protected, w.cost, dpl, ppl, bpl)
    ovec, oidx = oracle.interpret(cfg, table, pop, data, alpha=alpha, best_size=best_size)
    return vec, idx, ovec, oidx


def check_rank(vec, idx, ovec, oidx):
    # The chosen program must be one whose oracle error equals the oracle's best.
    assert np.float32(ovec[idx[0]]) == np.float32(ovec[oidx[0]]) or rel_close(ovec[idx[0]], ovec[oidx[0]], RTOL)


@pytest.mark.parametrize("strategy", [ppiamd.SEQ, ppiamd.PP, ppiamd.FP, ppiamd.PDP])
@pytest.mark.parametrize("form", [ppiamd.FORM_FP, ppiamd.FORM_PDP, ppiamd.FORM_PP])
def test_small_all_strategies_forms(eng, strategy, form):
    w = ppiamd.workload("small")
    print(f"case small strategy={strategy} form={form}", flush=True)
    vec, idx, ovec, oidx = run_both(eng, w, strategy, form)
    ok = rel_close(vec, ovec, RTOL)
    assert ok.all(), f"{(~ok).sum()} mismatches, first {np.nonzero(~ok)[0][:5]}: {vec[~ok][:5]} vs {ovec[~ok][:5]}"
    if form == ppiamd.FORM_PP and strategy in (ppiamd.SEQ, ppiamd.PP):
        # The PP form sums rows in the sequential path's order: bit-exact.
        assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32))
    check_rank(vec, idx, ovec, oidx)
    if np.array_equal(vec.view(np.uint32), ovec.view(np.uint32)):
        assert np.array_equal(idx, oidx)


@pytest.mark.parametrize("reducemax", [False, True])
@pytest.mark.parametrize("form", [ppiamd.FORM_FP, ppiamd.FORM_PDP])
@pytest.mark.parametrize("strategy", [ppiamd.SEQ, ppiamd.PDP])
def test_classification_bit_exact(eng, reducemax, form, strategy):
    w = ppiamd.workload("classification", n_programs=3000, nlin=20000, seed=7)
    print(f"case classification rm={reducemax} form={form} strategy={strategy}", flush=True)
    vec, idx, ovec, oidx = run_both(eng, w, strategy, form, reducemax=reducemax)
    assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32)), \
        f"{(vec != ovec).sum()} differ: {vec[vec != ovec][:5]} vs {ovec[vec != ovec][:5]}"
    assert np.array_equal(idx, oidx)


@pytest.mark.parametrize("rows_per_lane", [1, 2, 4, 8])
def test_regression_rows_per_lane(eng, rows_per_lane):
    w = ppiamd.workload("regression", n_programs=1500, nlin=5000, seed=3)
    print(f"case regression R={rows_per_lane}", flush=True)
    vec, idx, ovec, oidx = run_both(eng, w, ppiamd.SEQ, ppiamd.FORM_FP, rows_per_lane=rows_per_lane,
                                    programs_per_tile=32)
    ex, bad = oracle.exact_sums(oracle.make_cfg(ppiamd.SEQ, w.error_fn), w.grammar.table, w.pop, w.data)
    ok = sums_agree(vec, ovec, ex, bad, w.data.shape[0], ALPHA, w.pop.size, RTOL)
    assert ok.all(), f"{(~ok).sum()} mismatches: {vec[~ok][:5]} vs {ovec[~ok][:5]} (exact {ex[~ok][:5]})"
    check_rank(vec, idx, ovec, oidx)


@pytest.mark.parametrize("form", [ppiamd.FORM_FP, ppiamd.FORM_PDP, ppiamd.FORM_PP])
def test_full_operator_set(eng, form):
    """Every operator of interpreter_core on random programs, PROTECTED on and off."""
    gr = full_grammar(6)
    pop = ppiamd.grow(gr, 1200, 11, 7, 96, p_leaf=0.3)
    data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 12, 3000, 6) * np.float32(3.0)
    for prot in (True, False):
        w = ppiamd.Workload("full", gr, pop, data, ppiamd.EF_ABS, protected=prot)
        print(f"case full-ops form={form} protected={prot}", flush=True)
        vec, idx, ovec, oidx = run_both(eng, w, ppiamd.SEQ, form)
        ex, bad = oracle.exact_sums(oracle.make_cfg(ppiamd.SEQ, w.error_fn, protected=prot), gr.table, pop, data)
        ok = sums_agree(vec, ovec, ex, bad, data.shape[0], ALPHA, pop.size, RTOL)
        assert ok.all(), f"{(~ok).sum()} mismatches at {np.nonzero(~ok)[0][:5]}: {vec[~ok][:5]} vs {ovec[~ok][:5]}"


NON_NATIVE_BINARY = ["+", "-", "*", "mean", "max", "min", "mod", "pow", "xor", "=", "!=", "<", "<=", ">", ">=", "or",
                     "and"]
NON_NATIVE_UNARY = ["abs", "pow2", "pow3", "pow4", "pow5", "neg", "round", "ceil", "floor", "step", "sign", "gamma",
                    "not"]


@pytest.mark.parametrize("native", [False, True])
def test_operators_outside_the_native_set(eng, native):
    """functions.h:59-193 switch SIN, COS, EXP, EXP10, EXP2, DIV, SQRT, LOG, LOG10, LOG2 and TAN under
    NATIVE; POW, GAMMA (sinh, pow), MOD and the rest call the same routines in both builds
    (interpreter_core:44-132).  Programs over those operators alone give the sequential path's bits
    with NATIVE on: ordered form, SEQ semantics, bit for bit."""
    gr = ppiamd.Grammar(6, binary=NON_NATIVE_BINARY, unary=NON_NATIVE_UNARY)
    pop = ppiamd.grow(gr, 1500, 17, 7, 96, p_leaf=0.3)
    data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 18, 2500, 6) * np.float32(3.0)
    n_pow = int(np.isin(pop.phenotype, [gr.values["T_POW"], gr.values["T_GAMMA"]]).sum())
    assert n_pow > 500
    for prot in (True, False):
        w = ppiamd.Workload("non-native", gr, pop, data, ppiamd.EF_ABS, protected=prot)
        print(f"case non-native-ops native={native} protected={prot}", flush=True)
        vec, idx, ovec, oidx = run_both(eng, w, ppiamd.SEQ, ppiamd.FORM_ORDERED, native=native)
        same = vec.view(np.uint32) == ovec.view(np.uint32)
        assert same.all(), f"{(~same).sum()} differ at {np.nonzero(~same)[0][:5]}: {vec[~same][:5]} vs {ovec[~same][:5]}"
        assert list(idx) == list(oidx)


def test_edge_programs(eng):
    """Empty, invalid, unknown-symbol, out-of-range attribute, ternary and dead-code programs."""
    b = Builder(4)
    progs = [
        ["-", "x0", "x1"],                      # SUB x0 x1 = x0 - x1
        ["/", "x0", 0.0],                       # protected DIV -> 1
        ["log", 2.0],                           # protected LOG(2) = log(2)
        ["ite", ">", "x0", 0.0, "x1", "x2"],    # IF_THEN_ELSE
        ["ite", "x3", "+", "x0", "x1", "*", "x2", "x3"],
        ["x0", "x1", "x2"],                     # leftover stack entries: value is x0
        ["+", "x0"],                            # underflow: NaN
        [("sym", 4242)],                        # unknown symbol: NaN
        [("sym", b.t_attr, 9.0)],               # attribute column out of range: NaN
        [("sym", b.t_attr, 4.0)],               # column 4 = the target column itself
        ["exp", "exp", "exp", "x0"],            # overflows on some rows
        ["sqrt", "-", "x0", "x0"],              # sqrt(0)
        ["mod", "x0", 0.0],                     # fmod(x, 0) = NaN
        ["pow", "x0", 0.5],
        ["gamma", "abs", "x1"],
        ["tan", ("sym", b.values["T_CONST"], float(np.float32(1.57079637050629)))],
    ]
    pop = b.population(progs)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 5, 777, 4)
    empty = ppiamd.Population(pop.phenotype.copy(), pop.ephemeral.copy(), pop.size.copy(), pop.stride)
    empty.size[5] = 0
    for strategy in (ppiamd.SEQ, ppiamd.PP, ppiamd.FP, ppiamd.PDP):
        for form in (ppiamd.FORM_FP, ppiamd.FORM_PDP, ppiamd.FORM_PP):
            for p in (pop, empty):
                w = ppiamd.Workload("edge", None, p, data, ppiamd.EF_ABS)
                print(f"case edge strategy={strategy} form={form}", flush=True)
                vec, idx, ovec, oidx = run_both(eng, w, strategy, form, table=b.table)
                ok = rel_close(vec, ovec, RTOL)
                assert ok.all(), f"s={strategy} f={form}: {np.nonzero(~ok)[0]} {vec[~ok]} vs {ovec[~ok]}"


def test_prediction_mode(eng):
    b = Builder(4)
    prog = ["+", "*", "x0", "x0", "sin", "x3"]
    ph, ep = b.encode(prog)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 9, 1000, 4)
    eng.configure(b.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_ABS, compute_units=256)
    eng.set_dataset(data)
    out = eng.predict(np.array(ph, np.int32), np.array(ep, np.float32), len(ph), data.shape[0])
    pop = b.population([prog])
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS)
    ref, _ = oracle.interpret(cfg, b.table, pop, data, ppp_mode=1, prediction_mode=1)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("native", [False, True])
@pytest.mark.parametrize("prot", [True, False])
def test_constant_folding_same_bits(eng, prot, native):
    """Constant subtrees evaluated once at upload by fold_kernel give the same bits as the same
    programs evaluated with every subtree on every row (folding switched off), in both math modes."""
    gr = full_grammar(6)
    pop = ppiamd.grow(gr, 2000, 31, 8, 128, p_leaf=0.3)
    data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 13, 5000, 6) * np.float32(3.0)
    print(f"case folding protected={prot} native={native}", flush=True)
    res = {}
    for fold in (True, False):
        eng.configure(gr.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_ABS, protected=prot, native=native,
                      form=ppiamd.FORM_FP, fold_constants=fold)
        eng.set_dataset(data)
        vec, _ = eng.evaluate(pop, alpha=ALPHA, best_size=1)
        st = eng.stats()
        res[fold] = (vec.copy(), st.folded_subtrees, st.executed_nodes)
    assert res[True][1] > 100 and res[False][1] == 0
    assert res[True][2] < res[False][2]
    same = res[True][0].view(np.uint32) == res[False][0].view(np.uint32)
    assert same.all(), f"{(~same).sum()} programs differ, e.g. {np.nonzero(~same)[0][:5]}"


@pytest.mark.parametrize("nlin", [1, 2, 63, 65, 257, 1023])
def test_ragged_row_counts(eng, nlin):
    """Row counts that leave a single or partial row tile, with 37 programs and with one: every form
    against the oracle (PP and ordered bit-exact, FP and PDP within the float-sum bar)."""
    gr = full_grammar(4)
    data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 7, nlin, 4)
    for npro in (37, 1):
        pop = ppiamd.grow(gr, npro, 41 + nlin, 6, 48, p_leaf=0.3)
        dpl, ppl, bpl = oracle.reference_geometry(nlin, pop.n)
        cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS, False, True, None, dpl, ppl, bpl)
        ovec, oidx = oracle.interpret(cfg, gr.table, pop, data, alpha=ALPHA, best_size=1)
        ex, bad = oracle.exact_sums(oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS), gr.table, pop, data)
        for form in (ppiamd.FORM_FP, ppiamd.FORM_PDP, ppiamd.FORM_ORDERED, ppiamd.FORM_PP):
            print(f"case ragged nlin={nlin} programs={npro} form={form}", flush=True)
            eng.configure(gr.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_ABS, protected=True, form=form,
                          compute_units=256)
            eng.set_dataset(data)
            vec, idx = eng.evaluate(pop, alpha=ALPHA, best_size=1)
            if form in (ppiamd.FORM_PP, ppiamd.FORM_ORDERED):
                same = vec.view(np.uint32) == ovec.view(np.uint32)
                assert same.all(), f"form {form}: programs {np.nonzero(~same)[0][:5]}"
                assert idx[0] == oidx[0]
            else:
                ok = sums_agree(vec, ovec, ex, bad, nlin, ALPHA, pop.size)
                assert ok.all(), f"form {form}: {vec[~ok][:5]} vs {ovec[~ok][:5]}"
                check_rank(vec, idx, ovec, oidx)


@pytest.mark.parametrize("nfeat", [24, 40])
def test_wide_datasets(eng, nfeat):
    """Datasets wider than the 16 columns the FP form stages in LDS (the rest are read through the
    caches) and wider than 32: FP, ordered and PDP forms against the oracle."""
    gr = full_grammar(nfeat)
    pop = ppiamd.grow(gr, 600, 51 + nfeat, 7, 64, p_leaf=0.35)
    data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 8, 3000, nfeat) * np.float32(2.0)
    dpl, ppl, bpl = oracle.reference_geometry(data.shape[0], pop.n)
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_SQUARED, False, True, None, dpl, ppl, bpl)
    ovec, oidx = oracle.interpret(cfg, gr.table, pop, data, alpha=ALPHA, best_size=1)
    ex, bad = oracle.exact_sums(oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_SQUARED), gr.table, pop, data)
    for form in (ppiamd.FORM_FP, ppiamd.FORM_ORDERED, ppiamd.FORM_PDP):
        print(f"case wide nfeat={nfeat} form={form}", flush=True)
        eng.configure(gr.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_SQUARED, protected=True, form=form,
                      compute_units=256)
        eng.set_dataset(data)
        vec, idx = eng.evaluate(pop, alpha=ALPHA, best_size=1)
        if form == ppiamd.FORM_ORDERED:
            same = vec.view(np.uint32) == ovec.view(np.uint32)
            assert same.all(), f"ordered: programs {np.nonzero(~same)[0][:5]}"
        else:
            ok = sums_agree(vec, ovec, ex, bad, data.shape[0], ALPHA, pop.size)
            assert ok.all(), f"form {form}: {vec[~ok][:5]} vs {ovec[~ok][:5]}"
            check_rank(vec, idx, ovec, oidx)


@pytest.mark.parametrize("prot", [False, True])
def test_min_max_signed_zero(eng, prot):
    """MIN / MAX on equal operands return the second one, as the x86-64 C library's fmin / fmax
    (minsd / maxsd) do for the reference: min(-0, +0) = +0 and max(+0, -0) = -0.  Checked per row
    (prediction) on columns holding -0 and +0, and on constants (folded at upload)."""
    b = Builder(4)
    data = np.zeros((300, 5), dtype=np.float32)
    data[:, 0] = -0.0
    data[1::2, 0] = 0.0          # x0 alternates +0 / -0, x1 = +0, x2 = -0
    data[:, 2] = -0.0
    data[:, 3] = np.linspace(-2.0, 2.0, 300, dtype=np.float32)
    progs = [["min", "x0", "x1"], ["min", "x1", "x0"], ["max", "x0", "x2"], ["max", "x2", "x0"],
             ["min", "x2", "x1"], ["max", "x1", "x2"], ["gamma", "min", "x2", "x1"], ["gamma", "max", "x1", "x2"],
             ["min", "x3", "x3"], ["gamma", "min", "round", -0.25, "floor", 0.75],
             ["+", "x3", "gamma", "max", "floor", 0.75, "round", -0.25]]
    print(f"case signed-zero min/max protected={prot}", flush=True)
    eng.configure(b.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_ABS, protected=prot, compute_units=256)
    eng.set_dataset(data)
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS, protected=prot)
    for prog in progs:
        ph, ep = b.encode(prog)
        ph = np.array(ph, dtype=np.int32)
        ep = np.array(ep, dtype=np.float32)
        got = eng.predict(ph, ep, len(ph), data.shape[0])
        one = ppiamd.Population(ph, ep, np.array([len(ph)], dtype=np.int32), len(ph))
        want, _ = oracle.interpret(cfg, b.table, one, data, ppp_mode=1, prediction_mode=1)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"{prog}: {got[:4]} vs {want[:4]}"
    pop = b.population(progs)
    w = ppiamd.Workload("signed-zero", None, pop, data, ppiamd.EF_ABS, protected=prot)
    for form in (ppiamd.FORM_FP, ppiamd.FORM_PP):
        vec, idx, ovec, oidx = run_both(eng, w, ppiamd.SEQ, form, table=b.table)
        assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32)), f"form {form}: {vec} vs {ovec}"
