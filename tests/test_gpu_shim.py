"""GPU: the drop-in library (acc_*/seq_* signatures) driven like ppi.cc / ppp.cc, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import ppiamd
from ppiamd import _grammar as grammar
from helpers import rel_close, sums_agree
from test_boundary_cpu import SHIM, write_case

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ALPHA = np.float32(1e-6)


def shim_grammar():
    g = grammar.parse_bnf(open(os.path.join(ROOT, "tests", "shim", "grammar.bnf")).read())
    return g.engine_table(), g.values()


def population(n, seed, n_attr=4):
    table, vals = shim_grammar()
    gr = ppiamd.Grammar(n_attr)
    # Reuse the workload generator, then map the config grammar's symbols onto the shim grammar's.
    pop = ppiamd.grow(gr, n, seed, 7, 64)
    remap = {gr.values[k]: vals[k] for k in gr.values if k in vals}
    ph = pop.phenotype.copy()
    for a, b in remap.items():
        ph[pop.phenotype == a] = b
    pop.phenotype = ph
    return table, pop


def run_harness(tmp_path, variant, mode, data, pop, argv=(), best_size=1, alpha=ALPHA):
    inp, out = tmp_path / f"{mode}.in", tmp_path / f"{mode}.out"
    write_case(inp, data, pop, best_size, alpha)
    exe = os.path.join(SHIM, variant, "harness")
    print(f"harness {variant} {mode} {' '.join(argv)}", flush=True)
    r = subprocess.run([exe, mode, str(inp), str(out)] + list(argv), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.int32)
    cnt = raw[0]
    vec = raw[1:1 + cnt].view(np.float32)
    bs = raw[1 + cnt]
    idx = raw[2 + cnt:2 + cnt + bs]
    cb = raw[2 + cnt + bs:]
    return vec, idx, cb


@pytest.mark.parametrize("mode,strategy,argv", [
    ("acc", ppiamd.PDP, []),                      # reference default strategy
    ("acc", ppiamd.PP, ["-strategy", "PP"]),
    ("acc", ppiamd.FP, ["-strategy", "DP", "-v"]),
    ("acc", ppiamd.PDP, ["-strategy", "pdp", "-cl-mls", "128", "-acc-form", "PP"]),
    ("seq", ppiamd.SEQ, []),
])
def test_ge_generation(tmp_path, mode, strategy, argv):
    table, pop = population(800, 5)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 6, 1500, 4)
    vec, idx, cb = run_harness(tmp_path, "abs", mode, data, pop, argv, best_size=1)
    mls = 128 if "128" in argv else 0
    dpl, ppl, bpl = oracle.reference_geometry(data.shape[0], pop.n, mls=mls)
    cfg = oracle.make_cfg(strategy, ppiamd.EF_ABS, False, True, None, dpl, ppl, bpl)
    ovec, oidx = oracle.interpret(cfg, table, pop, data, alpha=ALPHA, best_size=1)
    ex, bad = oracle.exact_sums(oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS), table, pop, data)
    if strategy in (ppiamd.SEQ, ppiamd.PP) and "-acc-form" not in argv:
        # seq_* and -strategy PP default to the ordered form: the reference's row-order sums, bit for bit
        assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32)), \
            f"{int((vec.view(np.uint32) != ovec.view(np.uint32)).sum())} differ"
        assert list(idx) == list(oidx)
        ok = np.ones(pop.n, bool)
    elif strategy in (ppiamd.SEQ, ppiamd.PP):
        ok = sums_agree(vec, ovec, ex, bad, data.shape[0], ALPHA, pop.size)
    else:
        ok = rel_close(vec, ovec, 1e-5) | sums_agree(vec, ovec, ex, bad, data.shape[0], ALPHA, pop.size)
    assert ok.all(), f"{(~ok).sum()} differ"
    assert ovec[idx[0]] == ovec[oidx[0]] or rel_close(ovec[idx[0]], ovec[oidx[0]], 1e-5)
    if mode == "acc":
        assert list(cb) == [1, 1, 3]  # send + receive called once, nImmigrants from receive
    else:
        assert list(cb[:2]) == [0, 0]


@pytest.mark.parametrize("mode", ["seq", "seq-ppp", "seq-pred"])
def test_native_build_of_the_sequential_path(tmp_path, mode):
    """A -DNATIVE build of seq_interpret gives the same bits as a build without it: the sequential
    path's native_* shims are the double routines (sequential.cc:32-74).  Compared with the oracle
    (no NATIVE mode) and with the non-NATIVE build, bit for bit."""
    table, pop = population(1 if mode != "seq" else 700, 21)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 22, 1300, 4)
    vec, idx, _ = run_harness(tmp_path, "native", mode, data, pop, best_size=2)
    plain, pidx, _ = run_harness(tmp_path, "abs", mode, data, pop, best_size=2)
    assert np.array_equal(vec.view(np.uint32), plain.view(np.uint32))
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS)
    if mode == "seq":
        ovec, oidx = oracle.interpret(cfg, table, pop, data, alpha=ALPHA, best_size=2)
        assert list(idx) == list(oidx) == list(pidx)
    elif mode == "seq-ppp":
        ovec, _ = oracle.interpret(cfg, table, pop, data, alpha=0.0, ppp_mode=1)
    else:
        ovec, _ = oracle.interpret(cfg, table, pop, data, ppp_mode=1, prediction_mode=1)
    assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32))


def test_native_build_of_the_accelerator_path(tmp_path):
    """-DNATIVE on acc_*: the functions.h set (SIN, COS, EXP, LOG, DIV, ...) takes the hardware
    approximations, whose results are implementation-defined (OpenCL native_*), so only closeness is
    required; programs without those operators give the reference's bits."""
    table, pop = population(400, 23)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 24, 1000, 4)
    vec, idx, _ = run_harness(tmp_path, "native", "acc", data, pop, ["-strategy", "PP"])
    ovec, _ = oracle.interpret(oracle.make_cfg(ppiamd.PP, ppiamd.EF_ABS), table, pop, data, alpha=ALPHA)
    close = rel_close(vec, ovec, 1e-3)
    assert close.mean() > 0.9, close.mean()
    _, vals = shim_grammar()
    fast = {vals[k] for k in ("T_SIN", "T_COS", "T_EXP", "T_LOG", "T_DIV", "T_SQRT") if k in vals}
    plain = np.array([not (set(pop.phenotype[i * pop.stride:i * pop.stride + pop.size[i]].tolist()) & fast)
                      for i in range(pop.n)])
    assert plain.sum() > 10
    assert np.array_equal(vec[plain].view(np.uint32), ovec[plain].view(np.uint32))


@pytest.mark.parametrize("mode", ["acc", "seq"])
def test_classification_reducemax_cost(tmp_path, mode):
    """COST clamp + REDUCEMAX build: integer-exact error vector and identical ranking."""
    table, pop = population(600, 9)
    data = ppiamd.dataset(ppiamd.DS_CLUSTERS, 10, 2000, 4)
    vec, idx, cb = run_harness(tmp_path, "cost", mode, data, pop, ["-strategy", "PP"], best_size=3)
    strategy = ppiamd.PP if mode == "acc" else ppiamd.SEQ
    dpl, ppl, bpl = oracle.reference_geometry(data.shape[0], pop.n)
    cfg = oracle.make_cfg(strategy, ppiamd.EF_COST_CLAMP, True, True, ppiamd.COST4, dpl, ppl, bpl)
    ovec, oidx = oracle.interpret(cfg, table, pop, data, alpha=ALPHA, best_size=3)
    assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32))
    assert np.array_equal(idx, oidx)


@pytest.mark.parametrize("mode", ["acc-ppp", "seq-ppp", "acc-pred", "seq-pred"])
def test_single_solution(tmp_path, mode):
    """ppp.cc:98-140: one solution, error (alpha = 0) or per-row prediction."""
    table, pop = population(1, 13)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 14, 700, 4)
    vec, idx, cb = run_harness(tmp_path, "abs", mode, data, pop, ["-strategy", "PP"])
    cfg = oracle.make_cfg(ppiamd.SEQ if mode.startswith("seq") else ppiamd.PP, ppiamd.EF_ABS)
    if mode.endswith("pred"):
        ref, _ = oracle.interpret(cfg, table, pop, data, ppp_mode=1, prediction_mode=1)
        assert np.array_equal(vec.view(np.uint32), ref.view(np.uint32))
    else:
        ref, _ = oracle.interpret(cfg, table, pop, data, alpha=0.0, ppp_mode=1)
        ex, bad = oracle.exact_sums(oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS), table, pop, data)
        assert sums_agree(vec, ref, ex, bad, data.shape[0], 0.0, pop.size).all()


def test_profiling_variant_prints_reference_fields(tmp_path):
    table, pop = population(50, 2)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 3, 300, 4)
    inp, out = tmp_path / "p.in", tmp_path / "p.out"
    write_case(inp, data, pop)
    r = subprocess.run([os.path.join(SHIM, "abs", "harness_prof"), "acc", str(inp), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "time_kernel[1]:" in r.stdout and "gpops_kernel_communication:" in r.stdout


EXPR_EF = "(((X)-(Y))*((X)-(Y))/(1.0f+fabs(Y)))"  # tests/shim/expr/definitions.h


@pytest.mark.parametrize("mode,strategy,argv", [
    ("acc", ppiamd.PDP, []),
    ("acc", ppiamd.PP, ["-strategy", "PP"]),
    ("acc", ppiamd.FP, ["-strategy", "DP"]),
    ("seq", ppiamd.SEQ, []),
])
def test_user_error_function(tmp_path, mode, strategy, argv):
    """A build with a user -DEF (not one of the engine's built-ins): the drop-in hands the expression
    to the engine, which checks it against the build's own ERROR(X,Y) at init and evaluates every
    program in the reference's partition and row order for the strategy, so the error vector is the
    oracle's bit for bit (the oracle calling the same expression compiled by g++)."""
    from helpers import CompiledEF
    table, pop = population(600, 31)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 32, 1400, 4)
    vec, idx, _ = run_harness(tmp_path, "expr", mode, data, pop, argv, best_size=2)
    ref = CompiledEF([EXPR_EF], tmp_path)
    dpl, ppl, bpl = oracle.reference_geometry(data.shape[0], pop.n)
    cfg = oracle.make_cfg(strategy, ppiamd.EF_EXPR, False, True, None, dpl, ppl, bpl, ef_fn=ref.addr[EXPR_EF])
    ovec, oidx = oracle.interpret(cfg, table, pop, data, alpha=ALPHA, best_size=2)
    assert np.array_equal(vec.view(np.uint32), ovec.view(np.uint32)), \
        f"{int((vec.view(np.uint32) != ovec.view(np.uint32)).sum())} differ"
    assert list(idx) == list(oidx)
