"""GPU: the derived-column table follows the population (plan_derived's refill rule).  A full table
built from programs over columns 0-7 covers none of a population over columns 8-15, so the next
upload refills it from that population; its results are a fresh engine's, bit for bit, with the
same number of derived columns staged."""
import numpy as np
import pytest

import ppiamd
from helpers import Builder

pytestmark = pytest.mark.gpu
UN = ["sin", "cos", "exp", "log", "sqrt"]


def chains_population(b, rng, cols, n, chains=3, ops=5):
    progs = []
    for _ in range(n):
        prog = []
        for k in range(chains):
            if k < chains - 1:
                prog.append("+")
            prog += [UN[i] for i in rng.integers(0, len(UN), ops)] + [f"x{rng.choice(cols)}"]
        progs.append(prog)
    return b.population(progs)


def test_derived_table_refills_for_a_new_population():
    b = Builder(16)
    rng = np.random.default_rng(5)
    pa = chains_population(b, rng, list(range(0, 8)), 3000)
    pb = chains_population(b, rng, list(range(8, 16)), 1500)
    data = ppiamd.dataset(ppiamd.DS_POLY_NORMAL, 8, 20000, 16)
    e = ppiamd.Engine(0)
    e.configure(b.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_SQUARED, form=ppiamd.FORM_FP)
    e.set_dataset(data)
    e.evaluate(pa)
    full = e.stats().derived_columns
    assert full == 4096, full
    assert e.stats().derived_refills == 0
    vb, _ = e.evaluate(pb)
    refilled = e.stats().derived_columns
    assert e.stats().derived_refills == 1
    va, _ = e.evaluate(pa)  # and back: A is no longer covered either
    assert e.stats().derived_refills == 2
    e.close()
    f = ppiamd.Engine(0)
    f.configure(b.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_SQUARED, form=ppiamd.FORM_FP)
    f.set_dataset(data)
    fb, _ = f.evaluate(pb)
    fresh = f.stats().derived_columns
    f.close()
    g = ppiamd.Engine(0)
    g.configure(b.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_SQUARED, form=ppiamd.FORM_FP)
    g.set_dataset(data)
    ga, _ = g.evaluate(pa)
    g.close()
    assert refilled == fresh
    assert np.array_equal(vb.view(np.uint32), fb.view(np.uint32))
    assert np.array_equal(va.view(np.uint32), ga.view(np.uint32))
