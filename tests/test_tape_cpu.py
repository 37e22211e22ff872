"""CPU checks of the tape compiler: compiled tape executed on the host == oracle prefix-stack values."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
import ppiamd
from helpers import Builder, full_grammar

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def tape_lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("tape") / "libtape_interp.so"
    srcs = [os.path.join(HERE, "tape_interp.cc"), os.path.join(ROOT, "ppi-auto_amd", "csrc", "tape.cc")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", str(out)] + srcs,
                   check=True)
    L = C.CDLL(str(out))
    P = C.POINTER
    L.tape_eval.argtypes = [P(ppiamd.Symbol), C.c_int, P(C.c_int32), P(C.c_float), P(C.c_int32), C.c_int, C.c_int,
                            P(C.c_float), C.c_int, C.c_int, C.c_int, P(C.c_float), P(C.c_int32), P(C.c_int32)]
    L.tape_eval.restype = C.c_int
    L.tape_eval_ex.argtypes = L.tape_eval.argtypes + [C.c_int, P(C.c_int32)]
    L.tape_eval_ex.restype = C.c_int
    return L


def tape_values(L, table, pop, data, prot):
    data = np.ascontiguousarray(data, dtype=np.float32)
    out = np.zeros((pop.n, data.shape[0]), dtype=np.float32)
    slots = np.zeros(pop.n, dtype=np.int32)
    ncode = np.zeros(pop.n, dtype=np.int32)
    p = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    r = L.tape_eval(oracle._table(table), len(table), p(pop.phenotype, C.c_int32), p(pop.ephemeral, C.c_float),
                    p(pop.size, C.c_int32), pop.stride, pop.n, p(data, C.c_float), data.shape[0], data.shape[1],
                    int(prot), p(out, C.c_float), p(slots, C.c_int32), p(ncode, C.c_int32))
    assert r == 0
    return out, slots, ncode


def tape_values_ex(L, table, pop, data, prot, flags):
    """tape_eval_ex: flags bit 0 = derived columns, bit 1 = folded constant subtrees."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    out = np.zeros((pop.n, data.shape[0]), dtype=np.float32)
    slots = np.zeros(pop.n, dtype=np.int32)
    ncode = np.zeros(pop.n, dtype=np.int32)
    info = (C.c_int32 * 4)()
    p = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    r = L.tape_eval_ex(oracle._table(table), len(table), p(pop.phenotype, C.c_int32), p(pop.ephemeral, C.c_float),
                       p(pop.size, C.c_int32), pop.stride, pop.n, p(data, C.c_float), data.shape[0], data.shape[1],
                       int(prot), p(out, C.c_float), p(slots, C.c_int32), p(ncode, C.c_int32), flags, info)
    assert r == 0
    return out, slots, ncode, list(info)


def oracle_values(table, pop, data, prot):
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS, protected=prot)
    out = np.zeros((pop.n, data.shape[0]), dtype=np.float32)
    for i in range(pop.n):
        one = pop.subset(i + 1)
        single = ppiamd.Population(pop.phenotype[i * pop.stride:(i + 1) * pop.stride].copy(),
                                   pop.ephemeral[i * pop.stride:(i + 1) * pop.stride].copy(),
                                   pop.size[i:i + 1].copy(), pop.stride)
        del one
        out[i], _ = oracle.interpret(cfg, table, single, data, ppp_mode=1, prediction_mode=1)
    return out


def same_bits(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))


@pytest.mark.parametrize("prot", [True, False])
def test_tape_matches_prefix_stack_full_ops(tape_lib, prot):
    gr = full_grammar(5)
    pop = ppiamd.grow(gr, 300, 21, 8, 160, p_leaf=0.25)
    data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 4, 64, 5) * np.float32(4.0)
    tv, slots, ncode = tape_values(tape_lib, gr.table, pop, data, prot)
    ov = oracle_values(gr.table, pop, data, prot)
    ok = same_bits(tv, ov)
    assert ok.all(), f"{(~ok).sum()} values differ, programs {np.unique(np.nonzero(~ok)[0])[:10]}"
    # Sethi-Ullman bound: slots <= log2(leaves) + 1 for binary/unary trees.
    leaves = (pop.size + 1) // 2
    assert (slots <= np.floor(np.log2(np.maximum(leaves, 1))) + 1).all()
    # Folding terminal operands: never more instructions than symbols.
    assert (ncode <= pop.size).all()


def test_tape_edge_programs(tape_lib):
    b = Builder(4)
    progs = [["-", "x0", "x1"], ["-", 2.0, "x1"], ["-", "sin", "x0", "cos", "x1"], ["x0", "x1", "x2"],
             ["+", "x0"], [("sym", 4242)], [("sym", b.t_attr, 9.0)], [("sym", b.t_attr, 4.0)],
             ["ite", ">", "x0", 0.0, "x1", "x2"], ["ite", "x3", "+", "x0", "x1", "*", "x2", "x3"],
             ["ite", "ite", "x0", "x1", "x2", "ite", "x1", "x2", "x3", "-", "x3", "x0"],
             ["/", "*", "+", "x0", "x1", "-", "x2", "x3", "+", "*", "x0", "x2", "sin", "x1"]]
    pop = b.population(progs)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 5, 50, 4)
    tv, slots, ncode = tape_values(tape_lib, b.table, pop, data, True)
    ov = oracle_values(b.table, pop, data, True)
    ok = same_bits(tv, ov)
    assert ok.all(), f"programs {np.unique(np.nonzero(~ok)[0])}"
    assert np.isnan(tv[4]).all() and np.isnan(tv[5]).all() and np.isnan(tv[6]).all()
    assert np.array_equal(tv[7], data[:, 4])
    assert np.array_equal(tv[0], data[:, 0] - data[:, 1])


def test_deep_comb_program(tape_lib):
    """A 4,001-symbol left comb (+ + + ... x0 x1 ...) compiles without recursion and needs 1 slot."""
    b = Builder(4)
    n_ops = 2000
    prog = ["+"] * n_ops + [f"x{k % 4}" for k in range(n_ops + 1)]
    pop = b.population([prog])
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 5, 8, 4)
    tv, slots, ncode = tape_values(tape_lib, b.table, pop, data, True)
    ov = oracle_values(b.table, pop, data, True)
    assert same_bits(tv, ov).all()
    assert slots[0] <= 1 and ncode[0] == n_ops + 1


@pytest.mark.parametrize("prot", [True, False])
def test_derived_columns_match_prefix_stack(tape_lib, prot):
    """Chains of unary operators over input columns read precomputed derived columns: same values, bit for bit."""
    gr = full_grammar(5)
    pop = ppiamd.grow(gr, 300, 23, 8, 160, p_leaf=0.35)
    data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 4, 64, 5) * np.float32(4.0)
    data = np.ascontiguousarray(data, dtype=np.float32)
    out = np.zeros((pop.n, data.shape[0]), dtype=np.float32)
    slots = np.zeros(pop.n, dtype=np.int32)
    ncode = np.zeros(pop.n, dtype=np.int32)
    nder = (C.c_int32 * 4)()
    p = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    r = tape_lib.tape_eval_ex(oracle._table(gr.table), len(gr.table), p(pop.phenotype, C.c_int32),
                              p(pop.ephemeral, C.c_float), p(pop.size, C.c_int32), pop.stride, pop.n,
                              p(data, C.c_float), data.shape[0], data.shape[1], int(prot), p(out, C.c_float),
                              p(slots, C.c_int32), p(ncode, C.c_int32), 1, nder)
    assert r == 0
    assert nder[0] > 20, "the population must exercise derived columns"
    assert nder[1] > 5, "the population must exercise chained derived columns"
    _, _, ncode_plain = tape_values(tape_lib, gr.table, pop, data, prot)
    assert ncode.sum() < ncode_plain.sum()  # unary-of-column nodes cost no instruction
    ov = oracle_values(gr.table, pop, data, prot)
    ok = same_bits(out, ov)
    assert ok.all(), f"{(~ok).sum()} values differ, programs {np.unique(np.nonzero(~ok)[0])[:10]}"


@pytest.mark.parametrize("prot", [True, False])
def test_folded_constants_match_prefix_stack(tape_lib, prot):
    """Constant subtrees compiled into fold programs, evaluated once and written into the tape's
    placeholders (alone and with derived columns): the oracle's values, bit for bit."""
    gr = full_grammar(5)
    pop = ppiamd.grow(gr, 400, 29, 8, 160, p_leaf=0.3)
    data = ppiamd.dataset(ppiamd.DS_POLY_UNIFORM, 4, 48, 5) * np.float32(4.0)
    ov = oracle_values(gr.table, pop, data, prot)
    _, _, ncode_plain = tape_values(tape_lib, gr.table, pop, data, prot)
    for flags in (2, 3):
        tv, slots, ncode, info = tape_values_ex(tape_lib, gr.table, pop, data, prot, flags)
        ok = same_bits(tv, ov)
        assert ok.all(), f"flags {flags}: {(~ok).sum()} values differ, programs {np.unique(np.nonzero(~ok)[0])[:10]}"
        assert info[2] > 50, "the population must exercise folded subtrees"
        assert info[3] >= info[2]
        assert ncode.sum() < ncode_plain.sum()


def test_folded_edge_programs(tape_lib):
    """Whole-constant programs, nested and ternary constant subtrees, protected operators on constants,
    constants mixed with attributes, and folds feeding register-stack operands."""
    b = Builder(4)
    progs = [["+", 2.0, 3.0], ["sin", "+", 1.0, 2.0], ["ite", 1.0, 2.0, 3.0], ["ite", 0.0, 2.0, "x1"],
             ["*", "x0", "log", "+", 2.0, 3.0], ["log", -2.0], ["/", 1.0, 0.0], ["sqrt", -4.0],
             ["-", "exp", "cos", 0.5, "x2"], ["+", "*", "x0", "x1", "*", "sin", 1.5, "exp", 0.25],
             ["/", "*", "+", "x0", "x1", "-", 2.0, 3.0, "+", "*", 4.0, 5.0, "sin", "x1"],
             ["ite", ">", 1.0, 2.0, "x0", "sin", "*", 3.0, 0.5], ["exp", "exp", "exp", "exp", 3.0],
             ["+", "x0", 1.0], ["neg", "x3"], [2.5], ["mod", 7.0, 0.0], ["pow", -8.0, 0.5]]
    pop = b.population(progs)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 5, 40, 4)
    for prot in (True, False):
        ov = oracle_values(b.table, pop, data, prot)
        tv, slots, ncode, info = tape_values_ex(tape_lib, b.table, pop, data, prot, 2)
        ok = same_bits(tv, ov)
        assert ok.all(), f"protected={prot}: programs {np.unique(np.nonzero(~ok)[0])}"
        assert info[2] >= 13
        assert (ncode[[0, 1, 2, 5, 6, 7, 12, 16, 17]] == 1).all()  # one load of the folded value


def test_derived_table_follows_the_population(tape_lib):
    """A full derived-column table that covers under half of a new population's chain occurrences is
    refilled from that population (the same table a fresh start would give); one that covers it is
    kept; with refills off (the batched evaluator) it is never reset."""
    b = Builder(4)
    a_progs = [["sin", "x0"], ["cos", "x1"], ["exp", "x0"], ["+", "sin", "x0", "cos", "x1"], ["log", "exp", "x0"]] * 4
    b_progs = [["sin", "x2"], ["cos", "x3"], ["sqrt", "x2"], ["*", "sin", "x2", "cos", "x3"], ["exp", "sqrt", "x3"]] * 4
    pa, pb = b.population(a_progs), b.population(b_progs)
    L = tape_lib
    P = C.POINTER
    L.derived_refill.argtypes = [P(ppiamd.Symbol), C.c_int] + [P(C.c_int32), P(C.c_float), P(C.c_int32), C.c_int,
                                                                C.c_int] * 2 + [C.c_int, C.c_int, P(C.c_int32)]
    out = np.zeros(7, dtype=np.int32)
    p = lambda a, t: a.ctypes.data_as(P(t))  # noqa: E731
    args = []
    for pop in (pa, pb):
        args += [p(pop.phenotype, C.c_int32), p(pop.ephemeral, C.c_float), p(pop.size, C.c_int32), pop.stride, pop.n]
    assert L.derived_refill(oracle._table(b.table), len(b.table), *args, 4, 4, p(out, C.c_int32)) == 0
    cols_a, ep_a, ep_aa, ep_b, cols_b, same_as_fresh, ep_norefill = out
    assert cols_a == 4 and ep_a == 0          # filled to the cap, no reset
    assert ep_aa == 0                          # the same population is covered: kept
    assert ep_b == 1 and cols_b == 4           # B's chains are not covered: refilled from B
    assert same_as_fresh == 1
    assert ep_norefill == 0
