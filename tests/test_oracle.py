"""Pins the CPU oracle (oracle/ppi_oracle.c) before it is used to judge the engine.

Each check is chosen so that a plausible restatement mistake (operand order,
a dropped protection, a wrong reduction or tie rule, a wrong epilogue constant)
fails it: recorded outputs of the compiled reference, closed forms, an
independent recursive evaluator, hand-derived cases of the reference's
reduction trees, and finite differences for the FP64 gradient.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
import ppiamd
from helpers import Builder

HERE = os.path.dirname(os.path.abspath(__file__))
FLT_MAX = np.float32(3.4028234663852886e38)
F32 = np.float32


def pop_of(b, progs):
    return b.population(progs)


# --- recorded outputs of the compiled reference ------------------------------------------

def test_survey_known_answers():
    ka = json.load(open(os.path.join(HERE, "golden", "survey_known_answers.json")))
    b = Builder(4)
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS, protected=True)
    for case in ka["cases"]:
        ph, ep = b.encode(case["program"])
        if "value" in case:
            v = oracle.eval_row(cfg, b.table, ph, ep, np.array(case["row"], np.float32))
            assert F32(v) == F32(case["value"]), case["what"]
        else:
            data = np.array(case["rows"], np.float32)
            vec, _ = oracle.interpret(cfg, b.table, b.population([case["program"]]), data, alpha=case["alpha"])
            assert vec[0] == F32(case["fitness"]), case["what"]


# --- closed forms -------------------------------------------------------------------------

def seq_sum_f32(errs):
    return np.add.accumulate(np.asarray(errs, np.float32), dtype=np.float32)[-1]


@pytest.mark.parametrize("ef", [ppiamd.EF_ABS, ppiamd.EF_SQUARED])
def test_attribute_program_is_sequential_mean_error(ef):
    b = Builder(4)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 3, 777, 4)
    d = (data[:, 0] - data[:, 4]).astype(np.float32)
    errs = np.abs(d) if ef == ppiamd.EF_ABS else d * d
    want = F32(seq_sum_f32(errs) / F32(777)) + F32(1e-6) * F32(1)
    cfg = oracle.make_cfg(ppiamd.SEQ, ef)
    vec, _ = oracle.interpret(cfg, b.table, b.population([["x0"]]), data, alpha=1e-6)
    assert vec[0] == want


def test_constant_program():
    b = Builder(4)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 4, 100, 4)
    errs = np.abs(F32(2.5) - data[:, 4])
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS)
    vec, _ = oracle.interpret(cfg, b.table, b.population([[2.5]]), data, alpha=0.0)
    assert vec[0] == F32(seq_sum_f32(errs) / F32(100))


def test_non_finite_and_empty_rules():
    """sequential.cc:135-139 (size 0 -> FLT_MAX, no alpha) and :173,187 (first non-finite error ->
    FLT_MAX/nlin + alpha*size)."""
    b = Builder(4)
    data = np.zeros((10, 5), np.float32)
    data[:, 0] = np.linspace(0, 9, 10)
    data[7, 0] = 100.0  # exp(exp(100)) overflows on this row only
    pop = b.population([["exp", "exp", "x0"], ["x0"]])
    pop.size[1] = 0
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS)
    vec, _ = oracle.interpret(cfg, b.table, pop, data, alpha=1e-6)
    assert vec[0] == FLT_MAX / F32(10) + F32(1e-6) * F32(3)
    assert vec[1] == FLT_MAX


def test_reducemax_with_cost_is_zero_or_one():
    b = Builder(4)
    data = ppiamd.dataset(ppiamd.DS_CLUSTERS, 5, 400, 4)
    pop = b.population([["x0"], [0.0], [1.0], [2.0], [3.0]])
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_COST_CLAMP, reducemax=True, cost=ppiamd.COST4)
    vec, _ = oracle.interpret(cfg, b.table, pop, data, alpha=0.0)
    assert set(np.unique(vec)) <= {F32(0.0), F32(1.0)}


def test_cost_counts_are_integers_and_match_numpy():
    b = Builder(4)
    data = ppiamd.dataset(ppiamd.DS_CLUSTERS, 6, 999, 4)
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_COST_CLAMP, cost=ppiamd.COST4)
    vec, _ = oracle.interpret(cfg, b.table, b.population([["+", "x0", "x1"]]), data, alpha=0.0)
    x = (data[:, 0] + data[:, 1]).astype(np.float32)
    cls = np.trunc(np.clip(x, 0, 3)).astype(int)
    wrong = int((cls != data[:, 4].astype(int)).sum())
    assert vec[0] == F32(F32(wrong) / F32(999))


# --- an independent evaluator ------------------------------------------------------------

def rec_eval(tokens, row):
    """Recursive prefix evaluation (no stack scan) in float32 with double-rounded transcendentals."""
    pos = [0]

    def f():
        t = tokens[pos[0]]
        pos[0] += 1
        if isinstance(t, float):
            return F32(t)
        if isinstance(t, str) and t.startswith("x"):
            return F32(row[int(t[1:])])
        if t in ("+", "-", "*", "/"):
            a = f()
            c = f()
            if t == "+": return F32(a + c)
            if t == "-": return F32(a - c)
            if t == "*": return F32(a * c)
            return F32(1.0) if c == 0 else F32(a / c)
        a = f()
        if t == "sin": return F32(math.sin(float(a)))
        if t == "cos": return F32(math.cos(float(a)))
        if t == "exp":
            try:
                return F32(math.exp(float(a)))
            except OverflowError:
                return F32(np.inf)
        if t == "log": return F32(1.0) if a < 1 else F32(math.log(float(a)))
        if t == "sqrt": return F32(1.0) if a < 0 else F32(math.sqrt(float(a)))
        raise ValueError(t)
    return f()


def random_prog(rng, depth):
    if depth == 0 or rng.random() < 0.3:
        return [f"x{rng.integers(4)}"] if rng.random() < 0.6 else [float(F32(rng.uniform(-5, 5)))]
    if rng.random() < 0.5:
        return [str(rng.choice(["+", "-", "*", "/"]))] + random_prog(rng, depth - 1) + random_prog(rng, depth - 1)
    return [str(rng.choice(["sin", "cos", "exp", "log", "sqrt"]))] + random_prog(rng, depth - 1)


def test_stack_scan_matches_recursive_evaluation():
    rng = np.random.default_rng(7)
    b = Builder(4)
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS, protected=True)
    rows = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 8, 16, 4) * F32(2.0)
    with np.errstate(all="ignore"):
        for _ in range(300):
            prog = random_prog(rng, 5)
            ph, ep = b.encode(prog)
            for row in rows[:4]:
                want = rec_eval(prog, row)
                got = F32(oracle.eval_row(cfg, b.table, ph, ep, row))
                assert (got == want) or (np.isnan(got) and np.isnan(want)), (prog, row, got, want)


def test_textbook_identities():
    b = Builder(4)
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS, protected=True)
    xs = np.linspace(-20, 20, 101, dtype=np.float32)
    ident = ["+", "*", "sin", "x0", "sin", "x0", "*", "cos", "x0", "cos", "x0"]
    ph, ep = b.encode(ident)
    for x in xs:
        v = oracle.eval_row(cfg, b.table, ph, ep, np.array([x, 0, 0, 0, 0], np.float32))
        assert abs(v - 1.0) <= 4 * np.finfo(np.float32).eps
    ph, ep = b.encode(["exp", "log", "x0"])
    for x in np.linspace(1, 60, 60, dtype=np.float32):
        v = oracle.eval_row(cfg, b.table, ph, ep, np.array([x, 0, 0, 0, 0], np.float32))
        assert abs(v - x) <= 4 * np.finfo(np.float32).eps * x


# --- ranking and reduction trees of the accelerator path ----------------------------------

def test_pick_n_best_ties_go_to_the_largest_index():
    """Util.h:158-173: priority_queue<pair<-error, index>> pops equal errors largest index first."""
    assert list(oracle.pick_n_best(4, [1.0, 0.0, 0.0, 2.0])) == [2, 1, 0, 3]
    assert list(oracle.pick_n_best(2, [0.0, -0.0, 5.0])) == [1, 0]  # -0.0 == 0.0
    assert list(oracle.pick_n_best(2, [3.0, 1.0, 2.0], indices=[10, 11, 12])) == [11, 12]


def test_best_individual_tree_tie_rule():
    """accelerator.cl:270-313 with local size 4: lanes 1 and 2 tie; step s=2 moves lane 2 to
    position 0, step s=1 keeps position 0 on a tie, so the group reports index 2."""
    b = Builder(4)
    data = np.zeros((4, 5), np.float32)
    progs = [[9.0], [1.0], [1.0], [5.0], [1.0]]  # errors 9, 1, 1, 5 | 1
    pop = b.population(progs)
    cfg = oracle.make_cfg(ppiamd.PP, ppiamd.EF_ABS, best_local=4)
    _, idx = oracle.interpret(cfg, b.table, pop, data, alpha=0.0, best_size=2)
    # group 0 -> index 2 (error 1), group 1 -> index 4 (error 1); PickNBest ties -> larger group first
    assert list(idx) == [4, 2]


def test_pdp_one_versus_two_bad_lanes():
    """accelerator.cl:236,265: a lane that meets a non-finite error holds MAXFLOAT; one such lane gives
    MAXFLOAT/nlin, two overflow the tree sum to inf -> MAXFLOAT."""
    b = Builder(4)
    data = np.zeros((8, 5), np.float32)
    data[:, 0] = 1.0
    pop = b.population([["/", 1.0, "log", "x0"]])  # protected LOG(1) = 0, then 1/0 -> 1 (protected)
    prog = [["exp", "exp", "x0"]]
    pop = b.population(prog)
    cfg = oracle.make_cfg(ppiamd.PDP, ppiamd.EF_ABS, pdp_local=4)
    one = data.copy(); one[5, 0] = 100.0           # row 5 -> lane 1
    two = data.copy(); two[5, 0] = 100.0; two[6, 0] = 100.0  # lanes 1 and 2
    v1, _ = oracle.interpret(cfg, b.table, pop, one, alpha=0.0)
    v2, _ = oracle.interpret(cfg, b.table, pop, two, alpha=0.0)
    assert v1[0] == FLT_MAX / F32(8)
    assert v2[0] == FLT_MAX
    seq = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS)
    s2, _ = oracle.interpret(seq, b.table, pop, two, alpha=0.0)
    assert s2[0] == FLT_MAX / F32(8)


def test_dp_empty_program_depends_on_group_count():
    """accelerator.cl:105-108 + accelerator.cc:780-802: an empty program's partials are MAXFLOAT per
    group; one group gives MAXFLOAT/nlin, two or more overflow to FLT_MAX."""
    b = Builder(4)
    data = np.zeros((8, 5), np.float32)
    pop = b.population([["x0"]])
    pop.size[0] = 0
    v1, _ = oracle.interpret(oracle.make_cfg(ppiamd.FP, ppiamd.EF_ABS, dp_local=8), b.table, pop, data)
    v2, _ = oracle.interpret(oracle.make_cfg(ppiamd.FP, ppiamd.EF_ABS, dp_local=4), b.table, pop, data)
    assert v1[0] == FLT_MAX / F32(8)
    assert v2[0] == FLT_MAX


def test_pp_equals_seq_values():
    """evaluate_pp sums rows in the sequential order; with alpha added on the host it reproduces
    seq_interpret's vector exactly (accelerator.cl:22-84, accelerator.cc:849)."""
    w = ppiamd.workload("small", n_programs=300, nlin=300)
    a, _ = oracle.interpret(oracle.make_cfg(ppiamd.SEQ, w.error_fn), w.grammar.table, w.pop, w.data, alpha=1e-6)
    c, _ = oracle.interpret(oracle.make_cfg(ppiamd.PP, w.error_fn), w.grammar.table, w.pop, w.data, alpha=1e-6)
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


# --- FP64 gradient -------------------------------------------------------------------------

def test_fp64_gradient_closed_form_and_finite_differences():
    b = Builder(4)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_ADJOINT, 3, 500, 5)[:, [0, 1, 2, 3, 5]]
    progs = [["*", 1.5, "x0"], ["sin", "*", 0.7, "x1"], ["+", "*", 2.0, "x0", "exp", "*", -0.3, "x1"],
             ["/", "x2", 0.0], ["log", "+", 3.0, "x0"]]
    pop = b.population(progs)
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_SQUARED)
    fit, g, nc = oracle.grad_f64(cfg, b.table, pop, data)
    x0, y = data[:, 0].astype(np.float64), data[:, 4].astype(np.float64)
    # closed form: d/dc mean((c*x0 - y)^2) = mean(2 (c x0 - y) x0)
    assert g[0, 0] == pytest.approx(np.mean(2 * (1.5 * x0 - y) * x0), rel=1e-12)
    # protected DIV by the constant 0 is the constant 1: zero gradient
    assert g[3, 0] == 0.0
    h = 1e-3
    for i in range(pop.n):
        if i == 3:
            continue  # the protection point itself: finite differences straddle the discontinuity
        for k in range(nc[i]):
            pos = [j for j in range(pop.size[i]) if pop.phenotype[i * pop.stride + j] == b.t_const][k]
            e64 = pop.ephemeral.astype(np.float64)
            ep, em = e64.copy(), e64.copy()
            ep[i * pop.stride + pos] += h
            em[i * pop.stride + pos] -= h
            fd = (oracle.fitness_f64(cfg, b.table, pop, data, ep)[i] - oracle.fitness_f64(cfg, b.table, pop, data, em)[i]) / (2 * h)
            assert g[i, k] == pytest.approx(fd, rel=1e-5, abs=1e-9), (i, k)
