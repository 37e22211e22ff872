"""ctypes binding of the CPU oracle (oracle/libppi_oracle.so) — test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

import ppiamd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "libppi_oracle.so")


class OracleCfg(C.Structure):
    _fields_ = [
        ("strategy", C.c_int32), ("error_fn", C.c_int32), ("reducemax", C.c_int32), ("protected_ops", C.c_int32),
        ("cost_n", C.c_int32), ("cost", C.POINTER(C.c_float)), ("dp_local", C.c_int32), ("pdp_local", C.c_int32),
        ("best_local", C.c_int32), ("ef_fn", C.c_void_p),
    ]


GeGrammar = ppiamd.GeGrammar
GeParams = ppiamd.GeParams

_o = None


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        L = C.CDLL(ORACLE_LIB)
        P = C.POINTER
        L.oracle_interpret.argtypes = [P(OracleCfg), P(ppiamd.Symbol), C.c_int32, P(C.c_int32), P(C.c_float),
                                       P(C.c_int32), C.c_int32, C.c_int32, P(C.c_float), C.c_int32, C.c_int32,
                                       C.c_float, P(C.c_float), P(C.c_int32), P(C.c_int32), C.c_int32, C.c_int32]
        L.oracle_interpret.restype = C.c_int
        L.oracle_exact_sums.argtypes = [P(OracleCfg), P(ppiamd.Symbol), C.c_int32, P(C.c_int32), P(C.c_float),
                                        P(C.c_int32), C.c_int32, C.c_int32, P(C.c_float), C.c_int32, C.c_int32,
                                        P(C.c_double), P(C.c_int32)]
        L.oracle_exact_sums.restype = C.c_int
        L.oracle_eval_row.argtypes = [P(OracleCfg), P(ppiamd.Symbol), C.c_int32, P(C.c_int32), P(C.c_float),
                                      C.c_int32, P(C.c_float), C.c_int32]
        L.oracle_eval_row.restype = C.c_float
        L.oracle_pick_n_best.argtypes = [C.c_int32, P(C.c_int32), C.c_int32, P(C.c_float), P(C.c_int32)]
        L.oracle_pick_n_best.restype = None
        L.oracle_fitness_f64.argtypes = [P(OracleCfg), P(ppiamd.Symbol), C.c_int32, P(C.c_int32), P(C.c_double),
                                         P(C.c_int32), C.c_int32, C.c_int32, P(C.c_float), C.c_int32, C.c_int32,
                                         P(C.c_double)]
        L.oracle_fitness_f64.restype = C.c_int
        L.oracle_grad_f64.argtypes = [P(OracleCfg), P(ppiamd.Symbol), C.c_int32, P(C.c_int32), P(C.c_double),
                                      P(C.c_int32), C.c_int32, C.c_int32, P(C.c_float), C.c_int32, C.c_int32,
                                      C.c_int32, P(C.c_double), P(C.c_double), P(C.c_int32)]
        L.oracle_grad_f64.restype = C.c_int
        L.oracle_grad_mag_f64.argtypes = [P(OracleCfg), P(ppiamd.Symbol), C.c_int32, P(C.c_int32), P(C.c_double),
                                          P(C.c_int32), C.c_int32, C.c_int32, P(C.c_float), C.c_int32, C.c_int32,
                                          C.c_int32, P(C.c_double), P(C.c_double), C.c_int32]
        L.oracle_grad_mag_f64.restype = C.c_int
        L.oracle_time_seq.argtypes = [P(OracleCfg), P(ppiamd.Symbol), C.c_int32, P(C.c_int32), P(C.c_float),
                                      P(C.c_int32), C.c_int32, C.c_int32, P(C.c_float), C.c_int32, C.c_int32,
                                      C.c_float, P(C.c_float), C.c_int32]
        L.oracle_time_seq.restype = C.c_double
        L.oracle_rng_draws.argtypes = [C.c_uint64, C.c_int32, P(C.c_uint64), P(C.c_double)]
        L.oracle_ge_effective.argtypes = [P(GeGrammar), P(GeParams), P(C.c_int32), P(C.c_int32)]
        L.oracle_ge_decode.argtypes = [P(GeGrammar), P(GeParams), P(C.c_uint8), C.c_int32, P(C.c_int32),
                                       P(C.c_float), P(C.c_int32)]
        L.oracle_ge_breed.argtypes = [P(GeGrammar), P(GeParams), P(C.c_uint8), P(C.c_float), P(C.c_uint8), C.c_int32]
        L.oracle_ge_run.argtypes = [P(OracleCfg), P(ppiamd.Symbol), C.c_int32, P(GeGrammar), P(GeParams),
                                    P(C.c_float), C.c_int32, C.c_int32, C.c_int32, P(C.c_uint8), P(C.c_float),
                                    P(C.c_float), P(C.c_uint8)]
        _o = L
    return _o


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _table(table):
    t = (ppiamd.Symbol * max(1, len(table)))()
    for k, (s, op, v) in enumerate(table):
        t[k].symbol, t[k].op, t[k].value = s, op, v
    return t


def reference_geometry(nlin, nprog, cu=256, max_wg=1024, lds=160 * 1024, mls=0):
    """Work-group sizes acc_interpret_init derives (accelerator.cc:272-355)."""
    import math
    mls1 = mls if mls > 0 else min(max_wg, lds // 4)
    mls2 = min(max_wg, lds // 8)
    dp_local = max(1, min(mls1, math.ceil(np.float32(nlin) / np.float32(cu))))
    pdp_local = nlin if nlin < mls1 else mls1
    best_local = max(1, min(mls2, math.ceil(np.float32(nprog) / np.float32(cu))))
    return dp_local, pdp_local, best_local


def make_cfg(strategy, error_fn, reducemax=False, protected=True, cost=None, dp_local=1, pdp_local=1, best_local=1,
             ef_fn=None):
    """ef_fn: for EF_EXPR, the address of a compiled C function float(float X, float Y)."""
    c = OracleCfg()
    if ef_fn is not None:
        c.ef_fn = ef_fn
    c.strategy, c.error_fn, c.reducemax, c.protected_ops = strategy, error_fn, int(reducemax), int(protected)
    if cost is not None:
        c._cost = np.ascontiguousarray(cost, dtype=np.float32)
        c.cost_n = c._cost.shape[0]
        c.cost = _p(c._cost, C.c_float)
    c.dp_local, c.pdp_local, c.best_local = dp_local, pdp_local, best_local
    return c


def interpret(cfg, table, pop, data, alpha=0.0, best_size=0, ppp_mode=0, prediction_mode=0):
    """Error vector (and ranking) of the reference path named by cfg.strategy."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    nlin, ncol = data.shape
    out = np.zeros(nlin if (ppp_mode and prediction_mode) else pop.n, dtype=np.float32)
    idx = np.zeros(max(1, best_size), dtype=np.int32)
    bs = C.c_int32(best_size)
    t = _table(table)
    r = lib().oracle_interpret(C.byref(cfg), t, len(table), _p(pop.phenotype, C.c_int32), _p(pop.ephemeral, C.c_float),
                               _p(pop.size, C.c_int32), pop.stride, pop.n, _p(data, C.c_float), nlin, ncol,
                               alpha, _p(out, C.c_float), _p(idx, C.c_int32) if best_size else None,
                               C.byref(bs) if best_size else None, ppp_mode, prediction_mode)
    assert r == 0
    return out, (idx[:bs.value] if best_size else None)


def exact_sums(cfg, table, pop, data):
    """(exactly summed per-row float errors, bad flag) per program; see oracle_exact_sums."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    sums = np.zeros(pop.n, dtype=np.float64)
    bad = np.zeros(pop.n, dtype=np.int32)
    r = lib().oracle_exact_sums(C.byref(cfg), _table(table), len(table), _p(pop.phenotype, C.c_int32),
                                _p(pop.ephemeral, C.c_float), _p(pop.size, C.c_int32), pop.stride, pop.n,
                                _p(data, C.c_float), data.shape[0], data.shape[1], _p(sums, C.c_double),
                                _p(bad, C.c_int32))
    assert r == 0
    return sums, bad


def eval_row(cfg, table, phenotype, ephemeral, row):
    ph = np.ascontiguousarray(phenotype, dtype=np.int32)
    ep = np.ascontiguousarray(ephemeral, dtype=np.float32)
    rw = np.ascontiguousarray(row, dtype=np.float32)
    return lib().oracle_eval_row(C.byref(cfg), _table(table), len(table), _p(ph, C.c_int32), _p(ep, C.c_float),
                                 len(ph), _p(rw, C.c_float), len(rw))


def pick_n_best(n, errors, indices=None):
    e = np.ascontiguousarray(errors, dtype=np.float32)
    out = np.zeros(max(1, n), dtype=np.int32)
    ind = np.ascontiguousarray(indices, dtype=np.int32) if indices is not None else None
    lib().oracle_pick_n_best(n, _p(out, C.c_int32), len(e), _p(e, C.c_float),
                             _p(ind, C.c_int32) if ind is not None else None)
    return out[:n]


def fitness_f64(cfg, table, pop, data, eph64=None):
    data = np.ascontiguousarray(data, dtype=np.float32)
    e64 = np.ascontiguousarray(pop.ephemeral.astype(np.float64) if eph64 is None else eph64, dtype=np.float64)
    out = np.zeros(pop.n, dtype=np.float64)
    r = lib().oracle_fitness_f64(C.byref(cfg), _table(table), len(table), _p(pop.phenotype, C.c_int32),
                                 _p(e64, C.c_double), _p(pop.size, C.c_int32), pop.stride, pop.n,
                                 _p(data, C.c_float), data.shape[0], data.shape[1], _p(out, C.c_double))
    assert r == 0
    return out


def grad_f64(cfg, table, pop, data, max_consts=8, eph64=None):
    data = np.ascontiguousarray(data, dtype=np.float32)
    e64 = np.ascontiguousarray(pop.ephemeral.astype(np.float64) if eph64 is None else eph64, dtype=np.float64)
    fit = np.zeros(pop.n, dtype=np.float64)
    g = np.zeros((pop.n, max_consts), dtype=np.float64)
    nc = np.zeros(pop.n, dtype=np.int32)
    r = lib().oracle_grad_f64(C.byref(cfg), _table(table), len(table), _p(pop.phenotype, C.c_int32),
                              _p(e64, C.c_double), _p(pop.size, C.c_int32), pop.stride, pop.n,
                              _p(data, C.c_float), data.shape[0], data.shape[1], max_consts,
                              _p(fit, C.c_double), _p(g, C.c_double), _p(nc, C.c_int32))
    assert r == 0
    return fit, g, nc


def grad_mag_f64(cfg, table, pop, data, max_consts=16, eph64=None, f32_values=False):
    """FP64 gradient and the L1 magnitude of its per-row terms (both / nlin), constants in position order.
    f32_values: derivative of the float computation (float operator values, FP64 tangents)."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    e64 = np.ascontiguousarray(pop.ephemeral.astype(np.float64) if eph64 is None else eph64, dtype=np.float64)
    g = np.zeros((pop.n, max_consts), dtype=np.float64)
    m = np.zeros((pop.n, max_consts), dtype=np.float64)
    r = lib().oracle_grad_mag_f64(C.byref(cfg), _table(table), len(table), _p(pop.phenotype, C.c_int32),
                                  _p(e64, C.c_double), _p(pop.size, C.c_int32), pop.stride, pop.n,
                                  _p(data, C.c_float), data.shape[0], data.shape[1], max_consts,
                                  _p(g, C.c_double), _p(m, C.c_double), int(f32_values))
    assert r == 0
    return g, m


def time_seq(cfg, table, pop, data, alpha=0.0, threads=1):
    data = np.ascontiguousarray(data, dtype=np.float32)
    out = np.zeros(pop.n, dtype=np.float32)
    t = lib().oracle_time_seq(C.byref(cfg), _table(table), len(table), _p(pop.phenotype, C.c_int32),
                              _p(pop.ephemeral, C.c_float), _p(pop.size, C.c_int32), pop.stride, pop.n,
                              _p(data, C.c_float), data.shape[0], data.shape[1], alpha, _p(out, C.c_float), threads)
    return t, out


def rng_draws(seed, n):
    ints = np.zeros(n, dtype=np.uint64)
    reals = np.zeros(n, dtype=np.float64)
    lib().oracle_rng_draws(seed, n, _p(ints, C.c_uint64), _p(reals, C.c_double))
    return ints, reals


def ge_effective(ge):
    bg, mx = C.c_int32(), C.c_int32()
    assert lib().oracle_ge_effective(C.byref(ge.grammar), C.byref(ge.params), C.byref(bg), C.byref(mx)) == 0
    return bg.value, mx.value


def ge_decode(ge, genomes):
    """genomes: n x nb uint8 (0/1) -> phenotype, ephemeral (n x stride), size (n), stride."""
    g = np.ascontiguousarray(genomes, dtype=np.uint8)
    n = g.shape[0]
    _, mx = ge_effective(ge)
    ph = np.zeros((n, mx), dtype=np.int32)
    ep = np.zeros((n, mx), dtype=np.float32)
    sz = np.zeros(n, dtype=np.int32)
    assert lib().oracle_ge_decode(C.byref(ge.grammar), C.byref(ge.params), _p(g, C.c_uint8), n, _p(ph, C.c_int32),
                                  _p(ep, C.c_float), _p(sz, C.c_int32)) == 0
    return ph, ep, sz, mx


def ge_breed(ge, parents, fitness, n_immigrants=0):
    par = np.ascontiguousarray(parents, dtype=np.uint8)
    fit = np.ascontiguousarray(fitness, dtype=np.float32)
    out = np.zeros_like(par)
    assert lib().oracle_ge_breed(C.byref(ge.grammar), C.byref(ge.params), _p(par, C.c_uint8), _p(fit, C.c_float),
                                 _p(out, C.c_uint8), n_immigrants) == 0
    return out


def ge_run(cfg, table, ge, data, generations):
    data = np.ascontiguousarray(data, dtype=np.float32)
    P, nb = ge.params.population_size, ge.params.number_of_bits
    genomes = np.zeros((P, nb), dtype=np.uint8)
    fit = np.zeros(P, dtype=np.float32)
    hist = np.zeros(generations + 1, dtype=np.float32)
    best = np.zeros(nb, dtype=np.uint8)
    r = lib().oracle_ge_run(C.byref(cfg), _table(table), len(table), C.byref(ge.grammar), C.byref(ge.params),
                            _p(data, C.c_float), data.shape[0], data.shape[1], generations, _p(genomes, C.c_uint8),
                            _p(fit, C.c_float), _p(hist, C.c_float), _p(best, C.c_uint8))
    assert r == 0
    return genomes, fit, hist, best
