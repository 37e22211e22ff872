"""GPU: ppi_engine_evaluate_noalloc (every device buffer in the library's static scratch array) against
ppi_engine_evaluate, through the C ABI: identical vectors and rankings, bit for bit, on two calls in a
row (the second with fewer programs over the same rows), with the resident program set untouched."""
import numpy as np
import pytest

import ppiamd

pytestmark = pytest.mark.gpu

ALPHA = np.float32(1e-6)


@pytest.fixture(scope="module")
def eng():
    if ppiamd.device_count() == 0:
        pytest.fail("no HIP device visible")
    e = ppiamd.Engine(0)
    yield e
    e.close()


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


CASES = [
    ("regression", ppiamd.SEQ, ppiamd.FORM_FP),
    ("regression", ppiamd.FP, ppiamd.FORM_FP),
    ("regression", ppiamd.PP, ppiamd.FORM_FP),
    ("regression", ppiamd.PDP, ppiamd.FORM_PDP),
    ("classification", ppiamd.SEQ, ppiamd.FORM_FP),
    ("classification", ppiamd.PDP, ppiamd.FORM_AUTO),
]


@pytest.mark.parametrize("name,strategy,form", CASES)
def test_noalloc_matches_evaluate(eng, name, strategy, form):
    w = ppiamd.workload(name, n_programs=3000, nlin=20011)
    print(f"case noalloc {name} strategy={strategy} form={form}", flush=True)
    eng.configure(w.grammar.table, strategy=strategy, error_fn=w.error_fn, reducemax=w.reducemax,
                  protected=w.protected, cost=w.cost, form=form)
    eng.set_dataset(w.data)
    small = w.pop.subset(1234)
    ref_big, ridx_big = eng.evaluate(w.pop, alpha=ALPHA, best_size=3)
    ref_small, ridx_small = eng.evaluate(small, alpha=ALPHA, best_size=3)
    # resident set is now `small`; the no-allocation calls must leave it as it is
    v1, i1 = eng.evaluate_noalloc(w.pop, alpha=ALPHA, best_size=3)
    v2, i2 = eng.evaluate_noalloc(small, alpha=ALPHA, best_size=3)
    assert same_bits(v1, ref_big) and list(i1) == list(ridx_big)
    assert same_bits(v2, ref_small) and list(i2) == list(ridx_small)
    eng.launch()
    v3, i3 = eng.wait(small.n, alpha=ALPHA, best_size=3)
    assert same_bits(v3, ref_small) and list(i3) == list(ridx_small)


def test_noalloc_rejects_what_exceeds_the_scratch(eng):
    """5,000 programs over 1,000,000 rows need 5,000 x 3,907 row tiles x 4 B of partials (78 MB), more
    than the 64 MiB array: rejected on the host with PPI_E_ARG before anything is launched."""
    b = ppiamd.Grammar(4)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 3, 1_000_000, 4)
    pop = ppiamd.grow(b, 5000, 4, 3, 16)
    eng.configure(b.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_ABS, form=ppiamd.FORM_FP)
    eng.set_dataset(data)
    print("case noalloc over the limit", flush=True)
    with pytest.raises(ppiamd.PpiError) as e:
        eng.evaluate_noalloc(pop, alpha=ALPHA, best_size=1)
    assert e.value.status == ppiamd.E_ARG and "scratch" in str(e.value)
    # within the limit on the same rows: 2,000 programs (31 MB of partials)
    sub = pop.subset(2000)
    ref, ridx = eng.evaluate(sub, alpha=ALPHA, best_size=1)
    v, i = eng.evaluate_noalloc(sub, alpha=ALPHA, best_size=1)
    assert same_bits(v, ref) and list(i) == list(ridx)


def test_noalloc_rejects_other_forms(eng):
    w = ppiamd.workload("small", n_programs=50, nlin=300)
    eng.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, form=ppiamd.FORM_ORDERED)
    eng.set_dataset(w.data)
    with pytest.raises(ppiamd.PpiError) as e:
        eng.evaluate_noalloc(w.pop)
    assert e.value.status == ppiamd.E_UNSUPPORTED
