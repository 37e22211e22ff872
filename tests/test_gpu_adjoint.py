"""GPU checks of the reverse-mode adjoint (ppi_adjoint_*): gradients against the oracle's
forward-mode gradients (float computation and FP64), central finite differences, and the
on-device descent.  Bars as in test_adjoint_cpu.py: 1e-4 * (|gv| + mv) against the
derivative of the float computation, 1e-3 * (|g64| + m64) against FP64 where the float
computation's own derivative is within half that bar of FP64.

Which constants FP64 can judge is a property of the programs, not of the kernel: on
20,000 N(0,1) rows about 17% of this generator's constants have a float-computation
derivative that departs from FP64 by more than 5e-4 (extreme rows through exp, division
and sin), and the oracle's forward-mode gradient is NaN where 0 * inf occurs in its
tangents.  The tests require the judged share to stay above 0.75 so the FP64 comparison
is never vacuous."""
import numpy as np
import pytest

import oracle
import ppiamd
from helpers import Builder
from test_adjoint_cpu import check_both, diag, oracle_layout

pytestmark = pytest.mark.gpu


def adjoint_for(w, nlin=None):
    a = ppiamd.Adjoint(0)
    a.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=w.protected)
    a.set_dataset(w.data if nlin is None else w.data[:nlin])
    return a


@pytest.fixture(scope="module")
def adj_case():
    w = ppiamd.workload("adjoint", n_programs=1500, nlin=20000, seed=3)
    a = adjoint_for(w)
    fit = a.forward(w.pop, alpha=1e-6)
    return w, a, fit


@pytest.mark.parametrize("rows", [1, 2])
@pytest.mark.parametrize("partials", [ppiamd.PARTIALS_RECOMPUTE, ppiamd.PARTIALS_STORE])
def test_gradient_matches_oracle(adj_case, partials, rows):
    w, a, fit = adj_case
    print(f"adjoint parity: 1500 programs x 20000 rows, partials={partials}, rows per lane {rows}", flush=True)
    a.set_rows_per_lane(rows)
    try:
        g = a.backward(partials=partials)
    finally:
        a.set_rows_per_lane(0)
    st = a.stats()
    assert st.partials == partials and st.rows_per_lane == rows and st.n_consts > 2000
    sentinel = fit >= 3.0e38 / 20000
    gm = g.reshape(w.pop.n, w.pop.stride)
    assert (gm[sentinel] == 0).all()
    fit_ok = np.where(sentinel, np.nan, fit)
    off_v, off_64, judged = check_both(g, fit_ok, w.grammar.table, w.pop, w.data, True, ppiamd.EF_SQUARED,
                                       w.grammar.t_const)
    if off_v.any() or off_64.any():
        Gv, Mv, _ = oracle_layout(w.grammar.table, w.pop, w.data, True, ppiamd.EF_SQUARED, w.grammar.t_const,
                                  f32_values=True)
        G, M, _ = oracle_layout(w.grammar.table, w.pop, w.data, True, ppiamd.EF_SQUARED, w.grammar.t_const)
        assert not off_v.any(), f"{off_v.sum()} off the float-computation bar: " + diag(off_v, g, Gv, Mv, w.pop.stride)
        assert not off_64.any(), f"{off_64.sum()} off FP64: " + diag(off_64, g, G, M, w.pop.stride)
    print(f"FP64 judges {judged:.3f} of the constants", flush=True)
    assert judged >= 0.75, f"FP64 judges only {judged:.3f} of the constants"
    # Only T_CONST positions carry gradients.
    not_const = w.pop.phenotype != w.grammar.t_const
    assert (g[not_const] == 0).all()


def test_stores_agree(adj_case):
    """Partials in registers (recomputation) and in LDS (storage) give the same gradients
    up to the order the row sums are formed in (the two kernels may split rows into
    different ranges)."""
    w, a, fit = adj_case
    a.set_rows_per_lane(1)
    try:
        g1 = a.backward(partials=ppiamd.PARTIALS_RECOMPUTE).astype(np.float64)
        seg1 = a.stats().max_segments
        g2 = a.backward(partials=ppiamd.PARTIALS_STORE).astype(np.float64)
    finally:
        a.set_rows_per_lane(0)
    Gv, Mv, _ = oracle_layout(w.grammar.table, w.pop, w.data, True, ppiamd.EF_SQUARED, w.grammar.t_const,
                              f32_values=True)
    assert seg1 >= 2, "the register store must recompute some programs"
    same_nonfinite = ~np.isfinite(g1) & ~np.isfinite(g2)
    mag = np.where(np.isfinite(Mv), Mv, 0.0)  # NaN where the oracle's tangents hit 0 * inf
    bad = ~(same_nonfinite | (np.abs(g1 - g2) <= 1e-4 * (np.abs(g1) + mag) + 1e-30))
    assert not bad.any(), diag(bad, g1, g2, Mv, w.pop.stride)


def test_upstream_gradient_scales(adj_case):
    w, a, fit = adj_case
    g = a.backward(partials=ppiamd.PARTIALS_RECOMPUTE)
    scale = np.ldexp(1.0, np.arange(w.pop.n) % 7 - 3).astype(np.float32)  # powers of two: exact products
    gs = a.backward(grad_fitness=scale, partials=ppiamd.PARTIALS_RECOMPUTE)
    assert np.array_equal(gs, g * np.repeat(scale, w.pop.stride), equal_nan=True)


def test_edge_programs():
    b = Builder(4)
    c = 1.25
    progs = [["*", c, "x0"], [c], ["x1"], [], ["+", "x0"], ["ite", ">", "x0", 0.5, 2.0, "x1"],
             ["/", "x0", 0.0], ["log", "*", c, "x2"], ["exp", "*", 90.0, "x0"], ["round", "*", c, "x1"]]
    pop = b.population(progs)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 9, 3000, 4)
    a = ppiamd.Adjoint(0)
    a.configure(b.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_SQUARED, protected=True)
    a.set_dataset(data)
    fit = a.forward(pop)
    for partials, rows in [(p, r) for p in (ppiamd.PARTIALS_RECOMPUTE, ppiamd.PARTIALS_STORE) for r in (1, 2)]:
        a.set_rows_per_lane(rows)
        g = a.backward(partials=partials).reshape(pop.n, pop.stride)
        assert a.stats().rows_per_lane == rows
        x0 = data[:, 0].astype(np.float64)
        y = data[:, 4].astype(np.float64)
        cf = np.float64(np.float32(c))
        assert abs(g[0, 1] - np.mean(2 * (cf * x0 - y) * x0)) <= 1e-5 * np.mean(np.abs(2 * (cf * x0 - y) * x0))
        assert abs(g[1, 0] - np.mean(2 * (cf - y))) <= 1e-5 * np.mean(np.abs(2 * (cf - y)))
        assert (g[2] == 0).all() and (g[3] == 0).all() and (g[4] == 0).all()  # no constant / empty / invalid
        assert g[5, 3] == 0.0  # condition constant
        assert g[6, 2] == 0.0  # protected division by the constant 0
        assert (g[8] == 0).all() and fit[8] >= 3.0e38 / 3000  # exp overflows on some rows: sentinel, gradient 0
        assert g[9, 2] == 0.0  # under ROUND
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_SQUARED, protected=True)
    v, _ = oracle.interpret(cfg, b.table, pop, data, alpha=0.0)
    sentinel = (v >= 3.0e38 / 3000) | (fit >= 3.0e38 / 3000)
    assert np.array_equal(fit[sentinel], v[sentinel])
    assert (np.abs(fit - v) <= 1e-5 * np.abs(v))[~sentinel].all()


def test_rows_per_lane_needs_small_programs():
    """Two rows per lane hold programs of at most 5 stack and 5 adjoint slots: a complete
    binary tree of depth 7 needs 6 stack slots (a node whose operands are both leaves needs
    none), so a request for two rows fails with
    PPI_E_UNSUPPORTED while the measured choice falls back to one row per lane, with the
    same gradients as an explicit one-row run."""
    b = Builder(4)

    def tree(d, k):
        if d == 0:
            return [1.0 + 0.01 * k] if k % 3 == 0 else [f"x{k % 4}"]
        return ["+"] + tree(d - 1, 2 * k) + ["*"] + tree(d - 1, 2 * k + 1) + [0.5]

    progs = [tree(7, 0), ["*", 1.5, "x0"]]
    pop = b.population(progs)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 9, 2000, 4)
    a = ppiamd.Adjoint(0)
    a.configure(b.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_SQUARED, protected=True)
    a.set_dataset(data)
    a.forward(pop)
    a.set_rows_per_lane(2)
    with pytest.raises(ppiamd.PpiError):
        a.backward(partials=ppiamd.PARTIALS_STORE)
    a.set_rows_per_lane(0)
    g0 = a.backward()
    st = a.stats()
    assert st.rows_per_lane == 1 and all(t == -2.0 for t in st.tune_ms[4:])
    a.set_rows_per_lane(1)
    g1 = a.backward(partials=st.partials)
    assert np.array_equal(g0, g1, equal_nan=True)
    assert (g0.reshape(pop.n, pop.stride) != 0).any(axis=1).all()
    a.close()


def fd_check(w, g, fit, want=60, per_program=3):
    """Central differences of the FP64 fitness (h = 1e-3) on a sample of constants, all
    perturbed copies evaluated in one oracle call.  Two bars, each on the constants where it
    is meaningful:

    * magnitude bar, |g - fd| <= 1e-3 (|fd| + M): where the difference quotient is accurate
      (it agrees with the FP64 analytic gradient to 1e-4 (|G| + M)) and the float
      computation's own derivative agrees with FP64 (as in check_both); M is the magnitude
      of the summed per-row terms;
    * relative bar, |g - fd| <= 1e-3 |fd|: where the quotient agrees with the FP64 gradient
      to 1e-5 relative and the per-row terms cancel by less than 100x (M <= 100 |G|), so a
      float sum of them can be held to a relative bar at all.

    Returns (checked, strict) counts; asserts every checked constant passes its bars."""
    nlin = w.data.shape[0]
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_SQUARED, protected=True)
    G, M, _ = oracle_layout(w.grammar.table, w.pop, w.data, True, ppiamd.EF_SQUARED, w.grammar.t_const)
    Gv, _, _ = oracle_layout(w.grammar.table, w.pop, w.data, True, ppiamd.EF_SQUARED, w.grammar.t_const,
                             f32_values=True)
    G, M, Gv = (x.reshape(w.pop.n, w.pop.stride) for x in (G, M, Gv))
    g = g.reshape(w.pop.n, w.pop.stride)
    st = w.pop.stride
    picks = []
    for i in range(w.pop.n):
        if not np.isfinite(fit[i]) or fit[i] >= 3.0e38 / nlin:
            continue
        pos = np.nonzero(w.pop.phenotype[i * st:i * st + w.pop.size[i]] == w.grammar.t_const)[0]
        for p in pos[:per_program]:
            if np.isfinite(G[i, p]) and np.isfinite(Gv[i, p]) and np.isfinite(M[i, p]) and M[i, p] > 0:
                picks.append((i, p))
        if len(picks) >= 4 * want:
            break
    h = 1e-3
    k = len(picks)
    ph = np.concatenate([w.pop.phenotype[i * st:(i + 1) * st] for i, _ in picks] * 2)
    sz = np.concatenate([w.pop.size[i:i + 1] for i, _ in picks] * 2)
    e64 = np.concatenate([w.pop.ephemeral[i * st:(i + 1) * st].astype(np.float64) for i, _ in picks] * 2)
    for j, (_, p) in enumerate(picks):
        e64[j * st + p] += h
        e64[(k + j) * st + p] -= h
    pert = ppiamd.Population(ph, e64.astype(np.float32), sz, st)
    f = oracle.fitness_f64(cfg, w.grammar.table, pert, w.data, eph64=e64)
    checked = strict = 0
    misses = []
    for j, (i, p) in enumerate(picks):
        fd = (f[j] - f[k + j]) / (2 * h)
        if not np.isfinite(fd):
            continue
        Gi, Mi, gi = G[i, p], M[i, p], float(g[i, p])
        if abs(fd - Gi) > 1e-4 * (abs(Gi) + Mi):
            continue  # curvature or a kink within h: the quotient is not a reference here
        if abs(Gv[i, p] - Gi) > 0.5e-3 * (abs(Gi) + Mi):
            continue  # the float computation's own derivative departs from FP64 here
        if checked >= want:
            break
        checked += 1
        ok = abs(gi - fd) <= 1e-3 * (abs(fd) + Mi)
        if abs(fd - Gi) <= 1e-5 * abs(Gi) and Mi <= 100 * abs(Gi):
            strict += 1
            ok = ok and abs(gi - fd) <= 1e-3 * abs(fd)
        if not ok:
            misses.append((i, p, gi, fd, float(Gi), float(Gv[i, p]), float(Mi)))
    assert not misses, f"{len(misses)} of {checked} gradients off the central differences: {misses}"
    return checked, strict


def test_gradient_against_finite_differences():
    w = ppiamd.workload("adjoint", n_programs=120, nlin=5000, seed=11)
    a = adjoint_for(w)
    fit = a.forward(w.pop)
    g = a.backward(partials=ppiamd.PARTIALS_AUTO)
    checked, strict = fd_check(w, g, fit)
    print(f"finite differences: {checked} constants, {strict} under the relative 1e-3 bar", flush=True)
    assert checked >= 40, f"only {checked} constants with a reliable difference quotient"
    assert strict >= 20, f"only {strict} constants under the relative bar"


def test_descent_on_device():
    w = ppiamd.workload("adjoint", n_programs=800, nlin=20000, seed=5)
    a = adjoint_for(w)
    fit0 = a.forward(w.pop, alpha=1e-6)
    eph, fit = a.descend(20, step=0.1)
    st = a.stats()
    assert st.descend_steps == 20 and st.descend_ms > 0
    finite = fit0 < 3.0e38 / 20000
    # Best-so-far acceptance: never worse than the start.  The plain evaluator's
    # fitness of the returned constants is what `fit` holds.
    assert (fit[finite] <= fit0[finite]).all()
    assert (fit[~finite] == fit0[~finite]).all()
    with_consts = np.array([(w.pop.phenotype[i * w.pop.stride:i * w.pop.stride + w.pop.size[i]]
                             == w.grammar.t_const).any() for i in range(w.pop.n)])
    improved = (fit < fit0) & finite & with_consts
    assert improved.sum() >= 0.5 * (finite & with_consts).sum()
    assert np.median((fit0 - fit)[improved] / fit0[improved]) > 1e-3
    # Only T_CONST values move.
    moved = eph != w.pop.ephemeral
    assert (w.pop.phenotype[moved] == w.grammar.t_const).all()
    e = ppiamd.Engine(0)
    e.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
    e.set_dataset(w.data)
    refined = ppiamd.Population(w.pop.phenotype, eph, w.pop.size, w.pop.stride)
    v, _ = e.evaluate(refined, alpha=1e-6)
    assert np.array_equal(v, fit)
    # A further backward works on the refined programs.
    g = a.backward()
    assert np.isfinite(g).mean() > 0.99
