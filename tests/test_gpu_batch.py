"""GPU: batched persistent evaluator against the plain engine and the oracle."""
import numpy as np
import pytest

import oracle
import ppiamd
from helpers import sums_agree

pytestmark = pytest.mark.gpu
ALPHA = np.float32(1e-6)


@pytest.fixture(scope="module")
def small():
    w = ppiamd.batched_workload(8, seed=11, p_range=(50, 1500), len_range=(8, 96), rows_range=(300, 20000))
    tables = {tuple(g.table) for g in w.grammars}
    assert len(tables) == 1  # every pair's grammar numbers its symbols the same way
    return w


def run_batch(w, rows_per_lane, direct=0, **kw):
    b = ppiamd.Batch(0)
    b.configure(w.grammars[0].table, strategy=kw.pop("strategy", ppiamd.SEQ), error_fn=w.error_fn,
                rows_per_lane=rows_per_lane, compute_units=256, **kw)
    ids = [b.add_population(b.add_dataset(d), p) for p, d in zip(w.pops, w.datas)]
    b.set_direct_threshold(direct)
    b.build()
    b.run()
    return b, ids


def test_batched_equals_plain_fp_form_bit_for_bit(small):
    """Same rows per lane => same row tiles and reduction order as the plain FP form."""
    w = small
    b, ids = run_batch(w, 4)
    st = b.stats()
    print(f"items={st.n_items} grid={st.grid} R={st.rows_per_lane}", flush=True)
    eng = ppiamd.Engine(0)
    for k, (p, d) in enumerate(zip(w.pops, w.datas)):
        vec, idx = b.results(ids[k], alpha=ALPHA, best_size=2)
        eng.configure(w.grammars[k].table, strategy=ppiamd.SEQ, error_fn=w.error_fn, form=ppiamd.FORM_FP,
                      rows_per_lane=4, compute_units=256)
        eng.set_dataset(d)
        pv, pidx = eng.evaluate(p, alpha=ALPHA, best_size=2)
        assert np.array_equal(vec.view(np.uint32), pv.view(np.uint32)), f"pair {k}"
        assert np.array_equal(idx, pidx)
    eng.close()
    b.close()


@pytest.mark.parametrize("strategy", [ppiamd.SEQ, ppiamd.PDP])
@pytest.mark.parametrize("direct", [1, 3000, -1])
def test_direct_pairs_equal_plain(small, direct, strategy):
    """Pairs run as their own FP launch (all: threshold 1; the larger ones: 3,000 items; none: -1)
    give the plain engine's FP-form results bit for bit, through the batch's own epilogue."""
    w = small
    b, ids = run_batch(w, 4, direct=direct, strategy=strategy)
    st = b.stats()
    items = [p.n * ((d.shape[0] + 255) // 256) for p, d in zip(w.pops, w.datas)]
    want = len(items) if direct == 1 else (0 if direct < 0 else sum(i >= direct for i in items))
    print(f"direct={st.direct_pairs} of {len(items)} items={items} launches={st.n_shapes}", flush=True)
    assert st.direct_pairs == want
    if direct == 3000:
        assert 0 < want < len(items)
    eng = ppiamd.Engine(0)
    for k, (p, d) in enumerate(zip(w.pops, w.datas)):
        vec, idx = b.results(ids[k], alpha=ALPHA, best_size=2)
        eng.configure(w.grammars[k].table, strategy=strategy, error_fn=w.error_fn, form=ppiamd.FORM_FP,
                      rows_per_lane=4, compute_units=256)
        eng.set_dataset(d)
        pv, pidx = eng.evaluate(p, alpha=ALPHA, best_size=2)
        assert np.array_equal(vec.view(np.uint32), pv.view(np.uint32)), f"pair {k}"
        assert np.array_equal(idx, pidx)
    # a second replay gives the same bits
    first = [b.results(i, alpha=ALPHA)[0] for i in ids]
    b.run()
    for k in range(len(ids)):
        again, _ = b.results(ids[k], alpha=ALPHA)
        assert np.array_equal(again.view(np.uint32), first[k].view(np.uint32))
    eng.close()
    b.close()


def test_autotuned_shapes_per_pair():
    """The autotuner keys pairs by program length and rows and may give them different shapes;
    every pair's results equal the plain FP form's at the pair's own rows per lane, bit for bit."""
    w = ppiamd.batched_workload(10, seed=23, p_range=(200, 3000), len_range=(8, 200),
                                rows_range=(500, 300000))
    # one pair of very short programs, so that at least two length classes occur
    w.pops[0] = ppiamd.exact(w.grammars[0], 2000, 77, 3, 9)
    b, ids = run_batch(w, 0)
    st = b.stats()
    shapes = [b.pair_shape(i) for i in ids]
    print(f"shapes={shapes} launches={st.n_shapes} keys={st.tuned_keys} tune_ms={list(st.tune_ms)}",
          flush=True)
    assert st.autotuned == 1 and st.tuned_keys >= 2 and st.n_shapes >= 1
    assert st.n_shapes == len(set(shapes))
    eng = ppiamd.Engine(0)
    for k, (p, d) in enumerate(zip(w.pops, w.datas)):
        vec, idx = b.results(ids[k], alpha=ALPHA, best_size=2)
        eng.configure(w.grammars[k].table, strategy=ppiamd.SEQ, error_fn=w.error_fn, form=ppiamd.FORM_FP,
                      rows_per_lane=shapes[k][0], compute_units=256)
        eng.set_dataset(d)
        pv, pidx = eng.evaluate(p, alpha=ALPHA, best_size=2)
        assert np.array_equal(vec.view(np.uint32), pv.view(np.uint32)), f"pair {k} shape {shapes[k]}"
        assert np.array_equal(idx, pidx)
    eng.close()
    b.close()


@pytest.mark.parametrize("strategy", [ppiamd.SEQ, ppiamd.PDP])
def test_batched_autotuned_against_oracle(small, strategy):
    w = small
    b, ids = run_batch(w, 0, strategy=strategy)
    st = b.stats()
    assert st.autotuned == 1 and st.rows_per_lane in (2, 4, 8)
    print(f"autotuned R={st.rows_per_lane} tune_ms={list(st.tune_ms)}", flush=True)
    for k in (0, 3, 7):
        p, d = w.pops[k], w.datas[k]
        vec, idx = b.results(ids[k], alpha=ALPHA, best_size=1)
        dpl, ppl, bpl = oracle.reference_geometry(d.shape[0], p.n)
        cfg = oracle.make_cfg(strategy, w.error_fn, False, True, None, dpl, ppl, bpl)
        ovec, oidx = oracle.interpret(cfg, w.grammars[k].table, p, d, alpha=ALPHA, best_size=1)
        ex, bad = oracle.exact_sums(oracle.make_cfg(ppiamd.SEQ, w.error_fn), w.grammars[k].table, p, d)
        ok = sums_agree(vec, ovec, ex, bad, d.shape[0], ALPHA, p.size)
        if strategy == ppiamd.PDP:
            from helpers import rel_close
            ok |= rel_close(vec, ovec, 1e-5)
        assert ok.all(), f"pair {k}: {(~ok).sum()} differ"
    b.close()


def test_batched_classification_bit_exact():
    gr = ppiamd.Grammar(4)
    pops = [ppiamd.grow(gr, n, 20 + k, 8, 96) for k, n in enumerate((700, 90, 1300))]
    datas = [ppiamd.dataset(ppiamd.DS_CLUSTERS, 30 + k, r, 4) for k, r in enumerate((5000, 777, 12000))]
    b = ppiamd.Batch(0)
    b.configure(gr.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_COST_CLAMP, cost=ppiamd.COST4, reducemax=False,
                compute_units=256)
    ids = [b.add_population(b.add_dataset(d), p) for p, d in zip(pops, datas)]
    b.build()
    for rep in range(3):  # graph replays, with host work between them, must all agree
        b.run()
        for k in range(3):
            vec, idx = b.results(ids[k], alpha=ALPHA, best_size=1)
            cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_COST_CLAMP, cost=ppiamd.COST4)
            ovec, oidx = oracle.interpret(cfg, gr.table, pops[k], datas[k], alpha=ALPHA, best_size=1)
            assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32)), (rep, k)
            assert np.array_equal(idx, oidx)
    # next generation: new populations on the same staged datasets
    b.clear_populations()
    pops2 = [ppiamd.grow(gr, 400, 70 + k, 6, 64) for k in range(3)]
    ids = [b.add_population(k, p) for k, p in enumerate(pops2)]
    b.build()
    b.run()
    for k in range(3):
        vec, _ = b.results(ids[k], alpha=ALPHA)
        cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_COST_CLAMP, cost=ppiamd.COST4)
        ovec, _ = oracle.interpret(cfg, gr.table, pops2[k], datas[k], alpha=ALPHA)
        assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32))
    b.close()


def test_batched_shared_dataset_incremental_build():
    """Populations compiled together at build() on a thread pool, several sharing one dataset (and
    its derived-column table), give the plain FP engine's bits at the same rows per lane; a population added after a build is
    compiled by the next build and appended; a size outside [0, stride] is rejected when added."""
    gr = ppiamd.Grammar(6)
    data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 41, 6000, 6)
    pops = [ppiamd.grow(gr, n, 80 + k, 8, 96) for k, n in enumerate((900, 300, 1500, 64))]
    b = ppiamd.Batch(0)
    b.configure(gr.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_SQUARED, rows_per_lane=4)
    d = b.add_dataset(data)
    ids = [b.add_population(d, p) for p in pops[:3]]
    b.build()
    b.run()
    first = [b.results(i)[0] for i in ids]
    ids.append(b.add_population(d, pops[3]))
    b.build()
    b.run()
    e = ppiamd.Engine(0)  # same rows per lane: the same row tiles, so the same partial sums
    e.configure(gr.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_SQUARED, rows_per_lane=4, form=ppiamd.FORM_FP)
    e.set_dataset(data)
    for k, p in enumerate(pops):
        v, _ = b.results(ids[k])
        pv, _ = e.evaluate(p)
        assert np.array_equal(v.view(np.uint32), pv.view(np.uint32)), k
        if k < 3:
            assert np.array_equal(v.view(np.uint32), first[k].view(np.uint32)), k
    e.close()
    bad = pops[0].subset(4)
    bad.size[2] = bad.stride + 1
    with pytest.raises(ppiamd.PpiError):
        b.add_population(d, bad)
    b.close()
