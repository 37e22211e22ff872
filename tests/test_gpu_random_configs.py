"""GPU parity over randomly drawn configurations.

Each seed draws a grammar (a random subset of the operator set), a dataset shape (1 to 40
features, 1 to 3,000 rows), a population (1 to 1,500 programs, 16 to 200 symbols), the
result semantics (SEQ, PP, FP or PDP strategy, PROTECTED on or off, REDUCEMAX on or off,
absolute or squared error, or a classification error: inequality or a cost matrix, whose
integer entries may be negative) and a kernel form, and compares the engine with the
oracle: bit for bit where the form keeps the reference's row order (PP form; ordered
form) and for classification errors (integer sums), and with the float-sum bar (within
1e-5 relative, or at least as close to the exactly summed row errors as the oracle's own
sum) elsewhere.  One seed in four also runs the configuration through the batched
evaluator, as two pairs (the dataset and its first rows).  From seed 1,160 on, every other
seed rounds its regression dataset to multiples of 0.5, so exact zeros and integers reach
the operators' edge cases (poles, signed zeros, zero divisors).  The seed, shape and choices
are printed before each case.
"""
import numpy as np
import pytest

import oracle
import ppiamd
from helpers import ALL_BINARY, ALL_UNARY, rel_close, sums_agree

pytestmark = pytest.mark.gpu

ALPHA = np.float32(1e-6)


@pytest.fixture(scope="module")
def eng():
    if ppiamd.device_count() == 0:
        pytest.fail("no HIP device visible")
    e = ppiamd.Engine(0)
    yield e
    e.close()


def draw(seed):
    rng = np.random.default_rng(1000 + seed)
    nfeat = int(rng.integers(1, 41))
    nlin = int(rng.choice([1, 7, 64, 255, 256, 257, 1000, 3000]))
    nprog = int(rng.choice([1, 5, 300, 1500]))
    binary = list(rng.choice(ALL_BINARY, size=int(rng.integers(1, len(ALL_BINARY) + 1)), replace=False))
    unary = list(rng.choice(ALL_UNARY, size=int(rng.integers(0, len(ALL_UNARY) + 1)), replace=False))
    strategy = int(rng.choice([ppiamd.SEQ, ppiamd.PP, ppiamd.FP, ppiamd.PDP]))
    forms = [ppiamd.FORM_AUTO, ppiamd.FORM_FP, ppiamd.FORM_PDP, ppiamd.FORM_PP]
    if strategy in (ppiamd.SEQ, ppiamd.PP):
        forms.append(ppiamd.FORM_ORDERED)
    c = dict(nfeat=nfeat, nlin=nlin, nprog=nprog, binary=binary, unary=unary, strategy=strategy,
             form=int(rng.choice(forms)), protected=bool(rng.integers(0, 2)), reducemax=bool(rng.integers(0, 2)),
             error_fn=int(rng.choice([ppiamd.EF_ABS, ppiamd.EF_SQUARED])),
             max_symbols=int(rng.choice([16, 64, 200])), p_leaf=float(rng.uniform(0.25, 0.5)), cost=None)
    if rng.integers(0, 10) >= 7:
        c["error_fn"] = int(rng.choice([ppiamd.EF_NEQ, ppiamd.EF_COST, ppiamd.EF_COST_CLAMP]))
        if c["error_fn"] != ppiamd.EF_NEQ:
            c["cost"] = rng.integers(-2, 4, size=(4, 4)).astype(np.float32)
    c["batched"] = bool(rng.integers(0, 4) == 0)
    return c


def classification(c):
    return c["error_fn"] in (ppiamd.EF_NEQ, ppiamd.EF_COST, ppiamd.EF_COST_CLAMP)


def check(c, vec, idx, ovec, oidx, ex, bad, nlin, size):
    in_order = c["strategy"] in (ppiamd.SEQ, ppiamd.PP) and c["form"] in (ppiamd.FORM_PP, ppiamd.FORM_ORDERED)
    if in_order or classification(c):
        same = vec.view(np.uint32) == ovec.view(np.uint32)
        assert same.all(), f"programs {np.nonzero(~same)[0][:5]}: {vec[~same][:5]} vs {ovec[~same][:5]}"
        if idx is not None:
            assert idx[0] == oidx[0]
        return
    if c["reducemax"]:
        ok = rel_close(vec, ovec, 1e-5)  # a max is order-independent: only the sentinel rules differ
    else:
        ok = sums_agree(vec, ovec, ex, bad, nlin, ALPHA, size)
    assert ok.all(), f"programs {np.nonzero(~ok)[0][:5]}: {vec[~ok][:5]} vs {ovec[~ok][:5]}"
    if idx is not None:
        best = ovec[oidx[0]]
        assert np.float32(ovec[idx[0]]) == np.float32(best) or rel_close(ovec[idx[0]], best, 1e-5)


def oracle_run(c, gr, pop, data, mls=0, cu=256):
    dpl, ppl, bpl = oracle.reference_geometry(data.shape[0], pop.n, cu=cu, mls=mls)
    cfg = oracle.make_cfg(c["strategy"], c["error_fn"], c["reducemax"], c["protected"], c["cost"], dpl, ppl, bpl)
    ovec, oidx = oracle.interpret(cfg, gr.table, pop, data, alpha=ALPHA, best_size=1)
    ecfg = oracle.make_cfg(ppiamd.SEQ, c["error_fn"], c["reducemax"], c["protected"], c["cost"])
    ex, bad = oracle.exact_sums(ecfg, gr.table, pop, data)
    return ovec, oidx, ex, bad


def run_case(eng, seed, c, mls=0, cu=256):
    gr = ppiamd.Grammar(c["nfeat"], binary=c["binary"], unary=c["unary"])
    pop = ppiamd.grow(gr, c["nprog"], 7 + seed, 9, c["max_symbols"], p_leaf=c["p_leaf"])
    if classification(c):
        data = ppiamd.dataset(ppiamd.DS_CLUSTERS, 3 + seed, c["nlin"], c["nfeat"])
    else:
        data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 3 + seed, c["nlin"], c["nfeat"]) * np.float32(2.0)
        if seed >= 1160 and seed % 2 == 0:  # quantized: exact zeros, integers and halves in every column
            data = np.ascontiguousarray(np.round(data * np.float32(2.0)) / np.float32(2.0), dtype=np.float32)
    eng.configure(gr.table, strategy=c["strategy"], error_fn=c["error_fn"], reducemax=c["reducemax"],
                  protected=c["protected"], cost=c["cost"], form=c["form"], compute_units=cu,
                  max_local_size=mls)
    eng.set_dataset(data)
    vec, idx = eng.evaluate(pop, alpha=ALPHA, best_size=1)
    ovec, oidx, ex, bad = oracle_run(c, gr, pop, data, mls, cu)
    check(c, vec, idx, ovec, oidx, ex, bad, data.shape[0], pop.size)
    if not c["batched"]:
        return
    # The batched evaluator has no ordered or PP form: its sums follow the FP form's order.
    cb = dict(c, form=ppiamd.FORM_FP)
    datas = [data, np.ascontiguousarray(data[: max(1, data.shape[0] // 3)])]
    print(f"case random seed={seed} batched", flush=True)
    b = ppiamd.Batch(0)
    try:
        b.configure(gr.table, strategy=c["strategy"], error_fn=c["error_fn"], reducemax=c["reducemax"],
                    protected=c["protected"], cost=c["cost"], compute_units=cu,
                    max_local_size=mls)
        ids = [b.add_population(b.add_dataset(d), pop) for d in datas]
        b.build()
        b.run()
        for k, d in enumerate(datas):
            bvec, _ = b.results(ids[k], alpha=ALPHA)
            ovec, oidx, ex, bad = oracle_run(c, gr, pop, d, mls, cu)
            check(cb, bvec, None, ovec, oidx, ex, bad, d.shape[0], pop.size)
    finally:
        b.close()


@pytest.mark.parametrize("seed", range(1660))
def test_random_configuration(eng, seed):
    c = draw(seed)
    print(f"case random seed={seed} {c}", flush=True)
    run_case(eng, seed, c)


@pytest.mark.parametrize("seed", range(6000, 6200))
def test_random_geometry(eng, seed):
    """DP and PDP strategies under the reference geometry of other devices: work-group limits of
    1 to 1,000 (PDP lanes, not powers of two included) and 1 to 256 compute units (DP group size)."""
    c = draw(seed)
    r = np.random.default_rng(90000 + seed)
    mls, cu = int(r.choice([1, 3, 7, 100, 256, 1000])), int(r.choice([1, 7, 104, 256]))
    c["strategy"] = [ppiamd.FP, ppiamd.PDP][seed % 2]
    if c["form"] == ppiamd.FORM_ORDERED:
        c["form"] = ppiamd.FORM_FP
    print(f"case random geometry seed={seed} mls={mls} cu={cu} {c}", flush=True)
    run_case(eng, seed, c, mls, cu)
