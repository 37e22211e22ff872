"""CPU checks of the adjoint tape compiler (ppi::compile_adjoint), no GPU involved.

A host interpreter of the adjoint tape (tests/adjoint_interp.cc) runs forward
passes with partial records and the reverse sweeps in float, as the device
kernel does.  Its gradients are compared with the oracle's forward-mode (dual
number) gradients, an independent method, two ways:

  * float computation: dual numbers whose values are the sequential path's
    float results and whose tangents are FP64 derivatives at those values — the
    exact derivative of what the float evaluators compute.  Bar: |g - gv| <=
    1e-4 * (|gv| + mv), where mv = sum over rows of |EF' * d pred / d c| / nlin
    is the magnitude of the terms the gradient sums (rounding of each term moves
    the sum by a small multiple of it, whatever the cancellation);
  * FP64: everything in double.  Bar 1e-3 * (|g64| + m64), checked for the
    constants whose float-computation derivative is itself within half that bar
    of FP64; elsewhere the float forward pass departs from FP64 (sin of a huge
    argument, pow or mod near a discontinuity) and no float gradient can match.
    The share of such constants is bounded.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
import ppiamd
from helpers import ALL_BINARY, ALL_UNARY, Builder

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
UNARY_NO_GAMMA = [u for u in ALL_UNARY if u != "gamma"]
# Operators with a derivative (the rest cut the path to the constants and are tested separately).
DIFF_BINARY = ["+", "-", "*", "/", "mean", "max", "min", "mod", "pow"]


@pytest.fixture(scope="module")
def adj_lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("adj") / "libadj_interp.so"
    srcs = [os.path.join(HERE, "adjoint_interp.cc"), os.path.join(ROOT, "ppi-auto_amd", "csrc", "tape.cc")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", str(out)] + srcs,
                   check=True)
    L = C.CDLL(str(out))
    P = C.POINTER
    L.adj_eval.argtypes = [P(ppiamd.Symbol), C.c_int, P(C.c_int32), P(C.c_float), P(C.c_int32), C.c_int, C.c_int,
                           P(C.c_float), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_float), P(C.c_float),
                           P(C.c_int32)]
    L.adj_eval.restype = C.c_int
    return L


def adj_eval(L, table, pop, data, prot, ef, B):
    data = np.ascontiguousarray(data, dtype=np.float32)
    grad = np.zeros(pop.n * pop.stride, dtype=np.float32)
    fit = np.zeros(pop.n, dtype=np.float32)
    info = np.zeros(4, dtype=np.int32)
    p = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    r = L.adj_eval(oracle._table(table), len(table), p(pop.phenotype, C.c_int32), p(pop.ephemeral, C.c_float),
                   p(pop.size, C.c_int32), pop.stride, pop.n, p(data, C.c_float), data.shape[0], data.shape[1],
                   int(prot), ef, B, p(grad, C.c_float), p(fit, C.c_float), p(info, C.c_int32))
    assert r == 0
    return grad, fit, info


def oracle_layout(table, pop, data, prot, ef, t_const, max_consts=16, f32_values=False):
    """Oracle gradient and term magnitude in the ephemeral layout; mask of comparable programs."""
    cfg = oracle.make_cfg(ppiamd.SEQ, ef, protected=prot)
    g, m = oracle.grad_mag_f64(cfg, table, pop, data, max_consts, f32_values=f32_values)
    G = np.zeros(pop.n * pop.stride)
    M = np.zeros(pop.n * pop.stride)
    ok = np.zeros(pop.n, dtype=bool)
    for i in range(pop.n):
        ph = pop.phenotype[i * pop.stride:i * pop.stride + pop.size[i]]
        pos = np.nonzero(ph == t_const)[0]
        ok[i] = len(pos) <= max_consts
        k = min(len(pos), max_consts)
        G[i * pop.stride + pos[:k]] = g[i, :k]
        M[i * pop.stride + pos[:k]] = m[i, :k]
    return G, M, ok


def check_both(grad, fit, table, pop, data, prot, ef, t_const):
    """Returns (off the float-computation bar, off the FP64 bar, share of constants FP64 can judge).
    Programs whose float fitness is a non-finite sentinel have gradient 0 by definition and are
    checked separately."""
    Gv, Mv, ok = oracle_layout(table, pop, data, prot, ef, t_const, f32_values=True)
    G, M, _ = oracle_layout(table, pop, data, prot, ef, t_const)
    g = grad.astype(np.float64)
    # Terms beyond float range make the float sum overflow: not representable, not compared.
    mask = (np.repeat(ok & np.isfinite(fit), pop.stride) & np.isfinite(Gv) & np.isfinite(Mv) & (M > 0)
            & (Mv < 1e37))
    off_v = mask & ~(np.abs(g - Gv) <= 1e-4 * (np.abs(Gv) + Mv))
    bar = 1e-3 * (np.abs(G) + M)
    judged = mask & np.isfinite(G) & (np.abs(Gv - G) <= 0.5 * bar)
    off_64 = judged & ~(np.abs(g - G) <= bar)
    return off_v, off_64, judged.sum() / max(1, mask.sum())


def diag(bad, g, ref, mag, stride, k=6):
    idx = np.nonzero(bad)[0][:k]
    return "; ".join(f"prog {i // stride} pos {i % stride}: {g[i]!r} vs {ref[i]!r} (m {mag[i]:.3g})" for i in idx)


@pytest.mark.parametrize("prot", [True, False])
@pytest.mark.parametrize("ef", [ppiamd.EF_SQUARED, ppiamd.EF_ABS])
def test_adjoint_tape_matches_fp64_gradient(adj_lib, prot, ef):
    gr = ppiamd.Grammar(4, binary=DIFF_BINARY, unary=UNARY_NO_GAMMA)
    pop = ppiamd.grow(gr, 400, 31, 7, 96, p_const=0.45, max_consts=16)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 7, 200, 4)
    grad, fit, info = adj_eval(adj_lib, gr.table, pop, data, prot, ef, 16)
    off_v, off_64, judged = check_both(grad, fit, gr.table, pop, data, prot, ef, gr.t_const)
    finite = np.isfinite(fit)
    assert info[3] > 200, "too few live constants exercised"
    assert not off_v.any(), f"{off_v.sum()} gradients off, programs {np.unique(np.nonzero(off_v)[0] // pop.stride)[:10]}"
    assert not off_64.any(), f"{off_64.sum()} gradients off FP64"
    assert judged >= 0.9
    # Constants of programs whose fitness is a non-finite sentinel get gradient 0.
    gm = grad.reshape(pop.n, pop.stride)
    assert (gm[~finite] == 0).all()


def test_adjoint_config_grammar(adj_lib):
    """The adjoint benchmark's grammar and target (8 attributes, <= 8 constants per program)."""
    gr = ppiamd.Grammar(8)
    pop = ppiamd.grow(gr, 500, 5, 10, 128, max_consts=8)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_ADJOINT, 6, 300, 8)
    grad, fit, info = adj_eval(adj_lib, gr.table, pop, data, True, ppiamd.EF_SQUARED, 16)
    off_v, off_64, judged = check_both(grad, fit, gr.table, pop, data, True, ppiamd.EF_SQUARED, gr.t_const)
    assert not off_v.any(), f"{off_v.sum()} gradients off"
    assert not off_64.any(), f"{off_64.sum()} gradients off FP64"
    assert judged >= 0.9
    assert np.count_nonzero(grad) > 500


def test_segmentation_does_not_change_gradients(adj_lib):
    """Two partials per row (many forward recomputations) == 64 (one pass), bit for bit."""
    gr = ppiamd.Grammar(4)
    pop = ppiamd.exact(gr, 200, 77, 120, 240, p_const=0.8)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 8, 64, 4)
    g2, f2, i2 = adj_eval(adj_lib, gr.table, pop, data, True, ppiamd.EF_SQUARED, 2)
    g64, f64, i64 = adj_eval(adj_lib, gr.table, pop, data, True, ppiamd.EF_SQUARED, 64)
    assert i2[1] >= 4 and i64[1] <= 2
    assert i2[0] > pop.n, "programs with more than 16 constants must be split into several jobs"
    assert np.array_equal(g2.view(np.uint32), g64.view(np.uint32))
    assert np.array_equal(f2.view(np.uint32), f64.view(np.uint32))


def test_closed_forms(adj_lib):
    """Gradients that follow from calculus alone, on hand-written programs."""
    b = Builder(4)
    c = 1.7
    progs = [
        ["*", c, "x0"],                          # d/dc mean(EF(c*x0)) = mean(EF' * x0)
        ["sin", c],                              # d/dc = EF'(sin c) * cos c
        ["/", "x0", 0.0],                        # protected DIV by the constant 0: the value is 1, slope 0
        ["log", 0.5],                            # protected LOG below 1: slope 0
        ["ite", ">", "x0", 0.25, 2.0, -3.0],     # condition constant: 0; branches: selector
        ["-", c, c],                             # two constants: d/dc1 = mean(EF'), d/dc2 = -mean(EF')
        [c],                                     # the program is the constant
        ["round", "*", c, "x1"],                 # under ROUND: 0
    ]
    pop = b.population(progs)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 9, 100, 4).astype(np.float32)
    x0, x1, y = data[:, 0].astype(np.float64), data[:, 1].astype(np.float64), data[:, 4].astype(np.float64)
    grad, fit, _ = adj_eval(adj_lib, b.table, pop, data, True, ppiamd.EF_SQUARED, 16)
    gm = grad.reshape(pop.n, pop.stride)
    cf = np.float64(np.float32(c))

    def close(a, e):
        return abs(a - e) <= 1e-4 * max(1.0, abs(e))

    assert close(gm[0, 1], np.mean(2 * (cf * x0 - y) * x0))
    assert close(gm[1, 1], np.mean(2 * (np.sin(cf) - y) * np.cos(cf)))
    assert gm[2, 2] == 0.0
    assert gm[3, 1] == 0.0
    sel = x0 > 0.25
    assert gm[4, 3] == 0.0
    assert close(gm[4, 4], np.mean(np.where(sel, 2 * (2.0 - y), 0.0)))
    assert close(gm[4, 5], np.mean(np.where(sel, 0.0, 2 * (-3.0 - y))))
    assert close(gm[5, 1], np.mean(2 * (0.0 - y))) and close(gm[5, 2], -np.mean(2 * (0.0 - y)))
    assert close(gm[6, 0], np.mean(2 * (cf - y)))
    assert gm[7, 2] == 0.0
