"""GPU parity at the benchmark configurations' full sizes, against the oracle on the host's cores.

* regression: 30,000 programs x 100,000 rows x 16 features (squared error, sum), FP form, sequential-path
  semantics: every program within a relative 1e-5 of the oracle, or at least as close to the exact sum of
  the same per-row float errors as the oracle's own row-order float sum (the count of programs that pass
  only that way is printed); the ordered form bit for bit;
* classification: 30,000 programs x 150,000 rows (clamped 0/1 COST), sum and REDUCEMAX: integer counts,
  bit for bit, and the same ranking;
* batched: pairs of 100,000-300,000 rows whose programs have 150-256 symbols (register stacks beyond
  4 slots), one persistent launch: bit-identical to the plain engine per pair at the same row tile, and
  the oracle's bar;
* adjoint: 256 programs x 1,000,000 rows: gradients against the oracle's float-computation and FP64
  forward-mode gradients; 64 programs x 1,000,000 rows against central differences of the FP64 fitness.
The oracle runs OpenMP over programs on up to 16 host threads."""
import os

import numpy as np
import pytest

import oracle
import ppiamd
from helpers import sums_agree
from test_adjoint_cpu import check_both, diag, oracle_layout

pytestmark = pytest.mark.gpu

ALPHA = np.float32(1e-6)
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def eng():
    if ppiamd.device_count() == 0:
        pytest.fail("no HIP device visible")
    e = ppiamd.Engine(0)
    yield e
    e.close()


def oracle_seq(w, reducemax=None):
    rm = w.reducemax if reducemax is None else reducemax
    cfg = oracle.make_cfg(ppiamd.SEQ, w.error_fn, rm, w.protected, w.cost)
    t, vec = oracle.time_seq(cfg, w.grammar.table, w.pop, w.data, alpha=ALPHA, threads=THREADS)
    assert t >= 0
    return vec


def test_regression_full_size(eng):
    w = ppiamd.workload("regression")
    print(f"case full regression 30000 x 100000, oracle on {THREADS} threads", flush=True)
    eng.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=w.protected,
                  form=ppiamd.FORM_FP)
    eng.set_dataset(w.data)
    vec, idx = eng.evaluate(w.pop, alpha=ALPHA, best_size=1)
    ovec = oracle_seq(w)
    ex, bad = oracle.exact_sums(oracle.make_cfg(ppiamd.SEQ, w.error_fn, protected=w.protected), w.grammar.table,
                                w.pop, w.data)
    ok = sums_agree(vec, ovec, ex, bad, w.data.shape[0], ALPHA, w.pop.size)
    assert ok.all(), f"{(~ok).sum()} programs off the bar, e.g. {np.nonzero(~ok)[0][:5]}"
    a, b = vec.astype(np.float64), ovec.astype(np.float64)
    within = (a == b) | (np.abs(a - b) <= 1e-5 * np.maximum(np.abs(a), np.abs(b)))
    print(f"regression: {int(within.sum())} of {w.pop.n} within 1e-5 of the oracle, "
          f"{int((~within).sum())} only through the closer-to-exact branch; "
          f"{int((vec.view(np.uint32) == ovec.view(np.uint32)).sum())} bit-identical", flush=True)
    oidx = oracle.pick_n_best(1, ovec)
    assert ovec[idx[0]] == ovec[oidx[0]] or abs(float(ovec[idx[0]]) - float(ovec[oidx[0]])) <= 1e-5 * abs(float(ovec[oidx[0]]))
    # the ordered form: the sequential path's sums, bit for bit
    eng.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=w.protected,
                  form=ppiamd.FORM_ORDERED)
    eng.set_dataset(w.data)
    ovec2, oidx2 = eng.evaluate(w.pop, alpha=ALPHA, best_size=1)
    assert np.array_equal(ovec2.view(np.uint32), ovec.view(np.uint32))
    assert list(oidx2) == list(oidx)


@pytest.mark.parametrize("reducemax", [False, True])
def test_classification_full_size(eng, reducemax):
    w = ppiamd.workload("classification_max" if reducemax else "classification")
    print(f"case full classification 30000 x 150000 reducemax={reducemax}", flush=True)
    eng.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, reducemax=reducemax,
                  protected=w.protected, cost=w.cost, form=ppiamd.FORM_FP)
    eng.set_dataset(w.data)
    vec, idx = eng.evaluate(w.pop, alpha=ALPHA, best_size=3)
    ovec = oracle_seq(w, reducemax)
    same = vec.view(np.uint32) == ovec.view(np.uint32)
    assert same.all(), f"{(~same).sum()} differ, e.g. {np.nonzero(~same)[0][:5]}"
    assert list(idx) == list(oracle.pick_n_best(3, ovec))


def balanced(gr, depth, rng):
    """A complete binary tree of + - * over attribute leaves, in prefix order: 2^depth leaves whose
    Sethi-Ullman register need grows with depth (random trees of the same length rarely need more
    than 4 register-stack slots)."""
    ops = [gr.values["T_ADD"], gr.values["T_SUB"], gr.values["T_MULT"]]
    ph, ep = [], []

    def rec(d):
        if d == 0:
            ph.append(gr.t_attribute)
            ep.append(float(rng.integers(gr.n_attributes)))
            return
        ph.append(int(rng.choice(ops)))
        ep.append(0.0)
        rec(d - 1)
        rec(d - 1)
    rec(depth)
    return ph, ep


def test_batched_long_programs_large_rows():
    rng = np.random.default_rng(8)
    grammars, pops, datas = [], [], []
    for k, (nlin, nfeat) in enumerate([(100_000, 8), (200_000, 20), (300_000, 32)]):
        gr = ppiamd.Grammar(nfeat)
        grammars.append(gr)
        pop = ppiamd.exact(gr, 240, 900 + k, 150, 256)
        for i in range(0, 240, 12):  # every 12th program: a complete tree of 64 or 128 leaves
            ph, ep = balanced(gr, 6 + (i // 12) % 2, rng)
            pop.phenotype[i * pop.stride:i * pop.stride + len(ph)] = ph
            pop.ephemeral[i * pop.stride:i * pop.stride + len(ep)] = ep
            pop.size[i] = len(ph)
        pops.append(pop)
        datas.append(ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 950 + k, nlin, nfeat))
    print("case full batched: 3 pairs, 100k-300k rows, programs of 150-256 symbols", flush=True)
    b = ppiamd.Batch(0)
    b.configure(grammars[0].table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_SQUARED, protected=True)
    ids = [b.add_population(b.add_dataset(d), p) for p, d in zip(pops, datas)]
    b.build()
    b.run()
    res = [b.results(i, alpha=ALPHA, best_size=1) for i in ids]
    shapes = [b.pair_shape(i) for i in ids]
    b.close()
    with ppiamd.Engine(0) as e:
        for gr, p, d, (vec, idx), (R, _, S) in zip(grammars, pops, datas, res, shapes):
            assert S > 4, "the pair must run with the 9-slot stack"
            # the same row tile (64 x R rows) as the pair's batched launch: the same float partials
            e.configure(gr.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_SQUARED, protected=True,
                        form=ppiamd.FORM_FP, rows_per_lane=R)
            e.set_dataset(d)
            pv, pidx = e.evaluate(p, alpha=ALPHA, best_size=1)
            assert e.stats().max_slots > 4, "the populations must need the 9-slot stack"
            assert np.array_equal(vec.view(np.uint32), pv.view(np.uint32))
            w = ppiamd.Workload("pair", gr, p, d, ppiamd.EF_SQUARED)
            ovec = oracle_seq(w)
            ex, bad = oracle.exact_sums(oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_SQUARED), gr.table, p, d)
            ok = sums_agree(vec, ovec, ex, bad, d.shape[0], ALPHA, p.size)
            assert ok.all(), f"{(~ok).sum()} programs off the bar"


def test_adjoint_full_rows():
    w = ppiamd.workload("adjoint", n_programs=256, nlin=1_000_000, seed=3)
    print("case full adjoint: 256 programs x 1,000,000 rows", flush=True)
    a = ppiamd.Adjoint(0)
    a.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=w.protected)
    a.set_dataset(w.data)
    fit = a.forward(w.pop, alpha=1e-6)
    g = a.backward()
    a.close()
    sentinel = fit >= 3.0e38 / w.data.shape[0]
    fit_ok = np.where(sentinel, np.nan, fit)
    off_v, off_64, judged = check_both(g, fit_ok, w.grammar.table, w.pop, w.data, True, ppiamd.EF_SQUARED,
                                       w.grammar.t_const)
    if off_v.any() or off_64.any():
        Gv, Mv, _ = oracle_layout(w.grammar.table, w.pop, w.data, True, ppiamd.EF_SQUARED, w.grammar.t_const,
                                  f32_values=True)
        assert not off_v.any(), f"{off_v.sum()} off the float-computation bar: " + diag(off_v, g, Gv, Mv,
                                                                                       w.pop.stride)
        assert not off_64.any(), f"{off_64.sum()} off FP64"
    print(f"adjoint 1M rows: FP64 judges {judged:.3f} of the constants", flush=True)
    assert judged >= 0.7


def test_adjoint_finite_differences_full_rows():
    """The adjoint workload's 1,000,000 rows: device gradients against central differences of
    the FP64 fitness, under the magnitude bar and, where the terms cancel by less than 100x,
    the relative bar 1e-3 |fd| (test_gpu_adjoint.fd_check)."""
    from test_gpu_adjoint import fd_check
    w = ppiamd.workload("adjoint", n_programs=64, nlin=1_000_000, seed=7)
    print("case full adjoint finite differences: 64 programs x 1,000,000 rows", flush=True)
    a = ppiamd.Adjoint(0)
    a.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=w.protected)
    a.set_dataset(w.data)
    fit = a.forward(w.pop)
    g = a.backward()
    a.close()
    checked, strict = fd_check(w, g, fit, want=40)
    print(f"adjoint 1M rows, finite differences: {checked} constants, {strict} under the relative bar",
          flush=True)
    assert checked >= 25 and strict >= 10
