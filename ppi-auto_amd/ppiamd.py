"""Python binding (ctypes) of the PPI evaluation engine's C ABI, plus workload builders.

The engine is the HIP library ``libppi_engine.so`` built in this directory
(include/ppi_engine.h is its interface).  Nothing here evaluates programs: every
evaluation goes through the HIP library, and loading fails loudly when the
library is missing.  The workload builders produce the seeded populations and
datasets of the benchmark configurations through ``libppi_workload.so``.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import os
import sys
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "shim"))
import grammar as _grammar  # noqa: E402

ENGINE_LIB = os.path.join(HERE, "libppi_engine.so")
WORKLOAD_LIB = os.path.join(HERE, "libppi_workload.so")

# ppi_status
OK, E_ARG, E_STATE, E_PROGRAM, E_DEVICE, E_NOMEM, E_UNSUPPORTED, E_NODEVICE = range(8)
# ppi_strategy
SEQ, PP, FP, PDP = 0, 1, 2, 3
STRATEGIES = {"SEQ": SEQ, "PP": PP, "FP": FP, "DP": FP, "PDP": PDP}
# ppi_form
FORM_AUTO, FORM_PP, FORM_FP, FORM_PDP, FORM_ORDERED = 0, 1, 2, 3, 4
# ppi_error_fn
EF_ABS, EF_SQUARED, EF_NEQ, EF_COST, EF_COST_CLAMP, EF_EXPR = range(6)

OPS = _grammar.OPS


class PpiError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"ppi status {status}: {msg}")
        self.status = status


class Symbol(C.Structure):
    _fields_ = [("symbol", C.c_int32), ("op", C.c_int32), ("value", C.c_float)]


class Config(C.Structure):
    _fields_ = [
        ("strategy", C.c_int32), ("form", C.c_int32), ("error_fn", C.c_int32), ("reducemax", C.c_int32),
        ("protected_ops", C.c_int32), ("native_math", C.c_int32), ("cost_n", C.c_int32),
        ("cost", C.POINTER(C.c_float)), ("max_local_size", C.c_int32), ("compute_units", C.c_int32),
        ("rows_per_lane", C.c_int32), ("programs_per_tile", C.c_int32), ("ordered_pass_rows", C.c_int32),
        ("no_constant_folding", C.c_int32), ("ef_expr", C.c_char_p),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("upload_ms", C.c_double), ("kernel_ms", C.c_double), ("main_kernel_ms", C.c_double),
        ("download_ms", C.c_double), ("node_evals", C.c_double), ("executed_nodes", C.c_double),
        ("form", C.c_int32), ("rows_per_lane", C.c_int32), ("programs_per_tile", C.c_int32),
        ("max_slots", C.c_int32), ("derived_columns", C.c_int32), ("folded_subtrees", C.c_int32),
        ("folded_nodes", C.c_int32), ("derived_refills", C.c_int32),
    ]


_lib = None
_wl = None

class BatchStats(C.Structure):
    _fields_ = [
        ("graph_ms", C.c_double), ("kernel_ms", C.c_double), ("node_evals", C.c_double), ("build_ms", C.c_double),
        ("n_pairs", C.c_int32), ("n_items", C.c_int32), ("grid", C.c_int32), ("rows_per_lane", C.c_int32),
        ("feature_slots", C.c_int32), ("autotuned", C.c_int32), ("tune_ms", C.c_double * 4),
        ("derived_columns", C.c_int32), ("n_shapes", C.c_int32), ("tuned_keys", C.c_int32),
        ("direct_pairs", C.c_int32),
    ]


class AdjointStats(C.Structure):
    _fields_ = [
        ("compile_ms", C.c_double), ("backward_ms", C.c_double), ("kernel_ms", C.c_double),
        ("descend_ms", C.c_double), ("node_evals", C.c_double), ("forward_ins", C.c_double),
        ("reverse_ins", C.c_double), ("n_jobs", C.c_int32), ("n_consts", C.c_int32),
        ("max_segments", C.c_int32), ("partials", C.c_int32), ("rows_per_lane", C.c_int32),
        ("ranges", C.c_int32), ("descend_steps", C.c_int32), ("descend_reverted", C.c_int32),
        ("feature_slots", C.c_int32), ("pad_", C.c_int32), ("tune_ms", C.c_double * 8),
    ]


PARTIALS_AUTO, PARTIALS_RECOMPUTE, PARTIALS_STORE = 0, 1, 2


class GeGrammar(C.Structure):
    _fields_ = [("n_nonterminals", C.c_int32), ("n_rules", C.POINTER(C.c_int32)), ("rule_len", C.POINTER(C.c_int32)),
                ("rule_symbols", C.POINTER(C.c_int32)), ("max_rule", C.c_int32), ("t_const", C.c_int32),
                ("t_attribute", C.c_int32)]


class GeParams(C.Structure):
    _fields_ = [("population_size", C.c_int32), ("number_of_bits", C.c_int32), ("bits_per_gene", C.c_int32),
                ("bits_per_constant", C.c_int32), ("min_constant", C.c_float), ("max_constant", C.c_float),
                ("crossover_rate", C.c_float), ("mutation_rate", C.c_float), ("tournament_size", C.c_int32),
                ("elitism", C.c_int32), ("max_phenotype", C.c_int32), ("threads", C.c_int32),
                ("seed", C.c_int64), ("alpha", C.c_float)]


class GeStats(C.Structure):
    _fields_ = [("generation", C.c_int32), ("bits_per_gene", C.c_int32), ("max_size_phenotype", C.c_int32),
                ("max_slots", C.c_int32), ("best_fitness", C.c_float), ("breed_ms", C.c_double),
                ("decode_ms", C.c_double), ("compile_ms", C.c_double), ("evaluate_ms", C.c_double),
                ("rank_ms", C.c_double), ("generation_ms", C.c_double), ("node_evals", C.c_double)]


class GeSetup:
    """ppi_ge_grammar + ppi_ge_params of a BNF grammar (shim/grammar.py numbering) and ppi_init's
    options (defaults of ppi.cc:200-237: -nb 2000 -bg 8 -bc 16 -min -10 -max 10 -cp 0.9 -mr 0.01
    -ts 3), with the arrays kept alive."""

    def __init__(self, g, population_size=1024, number_of_bits=2000, bits_per_gene=8, bits_per_constant=16,
                 min_constant=-10.0, max_constant=10.0, crossover_rate=0.9, mutation_rate=0.01, tournament_size=3,
                 elitism=False, max_phenotype=0, threads=1, seed=1, alpha=1e-6):
        nts, alts = g.rule_tables()
        vals = g.values()
        self.max_rule = g.max_rule
        self._n_rules = np.array([len(a) for a in alts], dtype=np.int32)
        rules = [r for a in alts for r in a]
        self._rule_len = np.array([len(r) for r in rules], dtype=np.int32)
        self._rule_sym = np.zeros(len(rules) * g.max_rule, dtype=np.int32)
        for k, r in enumerate(rules):
            self._rule_sym[k * g.max_rule:k * g.max_rule + len(r)] = r
        self.grammar = GeGrammar(len(nts), _ptr(self._n_rules, C.c_int32), _ptr(self._rule_len, C.c_int32),
                                 _ptr(self._rule_sym, C.c_int32), g.max_rule, vals.get("T_CONST", -1),
                                 vals.get("T_ATTRIBUTE", -1))
        self.params = GeParams(population_size, number_of_bits, bits_per_gene, bits_per_constant, min_constant,
                               max_constant, crossover_rate, mutation_rate, tournament_size, int(elitism),
                               max_phenotype, threads, seed, alpha)

ENGINE_SYMBOLS = [
    "ppi_abi_version", "ppi_last_error", "ppi_device_count", "ppi_engine_create", "ppi_engine_destroy",
    "ppi_engine_configure", "ppi_engine_set_dataset", "ppi_engine_set_dataset_rowmajor", "ppi_engine_upload",
    "ppi_engine_launch", "ppi_engine_wait", "ppi_engine_evaluate", "ppi_engine_predict", "ppi_engine_get_stats",
    "ppi_engine_evaluate_noalloc", "ppi_engine_evaluate_ranked", "ppi_engine_resident_tape",
    "ppi_ge_create", "ppi_ge_destroy", "ppi_ge_configure", "ppi_ge_initialize", "ppi_ge_evolve",
    "ppi_ge_get_population", "ppi_ge_set_immigrants", "ppi_ge_best", "ppi_ge_decoded", "ppi_ge_tape",
    "ppi_ge_get_stats",
    "ppi_engine_time", "ppi_batch_create", "ppi_batch_destroy", "ppi_batch_configure", "ppi_batch_add_dataset",
    "ppi_batch_add_population", "ppi_batch_clear_populations", "ppi_batch_build", "ppi_batch_launch",
    "ppi_batch_wait", "ppi_batch_results", "ppi_batch_time", "ppi_batch_get_stats", "ppi_batch_pair_shape",
    "ppi_batch_set_direct_threshold",
    "ppi_adjoint_create", "ppi_adjoint_destroy", "ppi_adjoint_configure", "ppi_adjoint_set_dataset_rowmajor",
    "ppi_adjoint_forward", "ppi_adjoint_backward", "ppi_adjoint_descend", "ppi_adjoint_time",
    "ppi_adjoint_get_stats", "ppi_adjoint_set_rows_per_lane", "ppi_ef_evaluate",
]


def lib() -> C.CDLL:
    """Load the HIP engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(ENGINE_LIB):
            raise RuntimeError(f"{ENGINE_LIB} is missing: run `make -C {HERE}` (or __graft_entry__.build())")
        L = C.CDLL(ENGINE_LIB)
        P = C.POINTER
        L.ppi_abi_version.restype = C.c_int32
        L.ppi_last_error.restype = C.c_char_p
        L.ppi_device_count.argtypes = [P(C.c_int32)]
        L.ppi_engine_create.argtypes = [P(C.c_void_p), C.c_int32]
        L.ppi_engine_destroy.argtypes = [C.c_void_p]
        L.ppi_engine_destroy.restype = None
        L.ppi_engine_configure.argtypes = [C.c_void_p, P(Config), P(Symbol), C.c_int32]
        L.ppi_engine_set_dataset.argtypes = [C.c_void_p, P(P(C.c_float)), C.c_int32, C.c_int32]
        L.ppi_engine_set_dataset_rowmajor.argtypes = [C.c_void_p, P(C.c_float), C.c_int32, C.c_int32]
        L.ppi_engine_upload.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_float), P(C.c_int32), C.c_int32, C.c_int32]
        L.ppi_engine_launch.argtypes = [C.c_void_p]
        L.ppi_engine_wait.argtypes = [C.c_void_p, C.c_float, P(C.c_float), P(C.c_int32), P(C.c_int32)]
        L.ppi_engine_evaluate.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_float), P(C.c_int32), C.c_int32,
                                          C.c_int32, C.c_float, P(C.c_float), P(C.c_int32), P(C.c_int32)]
        L.ppi_engine_predict.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_float), C.c_int32, P(C.c_float)]
        L.ppi_engine_evaluate_ranked.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_float), P(C.c_int32), C.c_int32,
                                                 C.c_int32, C.c_float, P(C.c_float), P(C.c_int32), P(C.c_int32)]
        L.ppi_engine_resident_tape.argtypes = [C.c_void_p, P(C.c_uint32), P(C.c_uint32), P(C.c_int64)]
        L.ppi_ge_create.argtypes = [P(C.c_void_p), C.c_void_p]
        L.ppi_ge_destroy.argtypes = [C.c_void_p]
        L.ppi_ge_destroy.restype = None
        L.ppi_ge_configure.argtypes = [C.c_void_p, P(GeParams), P(GeGrammar)]
        L.ppi_ge_initialize.argtypes = [C.c_void_p]
        L.ppi_ge_evolve.argtypes = [C.c_void_p, C.c_int32]
        L.ppi_ge_get_population.argtypes = [C.c_void_p, P(C.c_uint8), P(C.c_float)]
        L.ppi_ge_set_immigrants.argtypes = [C.c_void_p, P(C.c_uint8), C.c_int32]
        L.ppi_ge_best.argtypes = [C.c_void_p, P(C.c_float), P(C.c_uint8)]
        L.ppi_ge_decoded.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_float), P(C.c_int32), P(C.c_int32)]
        L.ppi_ge_tape.argtypes = [C.c_void_p, P(C.c_uint32), P(C.c_uint32), P(C.c_int64)]
        L.ppi_ge_get_stats.argtypes = [C.c_void_p, P(GeStats)]
        L.ppi_engine_evaluate_noalloc.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_float), P(C.c_int32), C.c_int32,
                                                  C.c_int32, C.c_float, P(C.c_float), P(C.c_int32), P(C.c_int32)]
        L.ppi_engine_get_stats.argtypes = [C.c_void_p, P(Stats)]
        L.ppi_engine_time.argtypes = [C.c_void_p, C.c_int32, P(C.c_double)]
        L.ppi_batch_create.argtypes = [P(C.c_void_p), C.c_int32]
        L.ppi_batch_destroy.argtypes = [C.c_void_p]
        L.ppi_batch_destroy.restype = None
        L.ppi_batch_configure.argtypes = [C.c_void_p, P(Config), P(Symbol), C.c_int32]
        L.ppi_batch_add_dataset.argtypes = [C.c_void_p, P(C.c_float), C.c_int32, C.c_int32, P(C.c_int32)]
        L.ppi_batch_add_population.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_float), P(C.c_int32),
                                               C.c_int32, C.c_int32, P(C.c_int32)]
        L.ppi_batch_clear_populations.argtypes = [C.c_void_p]
        L.ppi_batch_build.argtypes = [C.c_void_p]
        L.ppi_batch_launch.argtypes = [C.c_void_p]
        L.ppi_batch_wait.argtypes = [C.c_void_p]
        L.ppi_batch_results.argtypes = [C.c_void_p, C.c_int32, C.c_float, P(C.c_float), P(C.c_int32), P(C.c_int32)]
        L.ppi_batch_time.argtypes = [C.c_void_p, C.c_int32, P(C.c_double)]
        L.ppi_batch_get_stats.argtypes = [C.c_void_p, P(BatchStats)]
        L.ppi_batch_pair_shape.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_int32)]
        L.ppi_batch_set_direct_threshold.argtypes = [C.c_void_p, C.c_int64]
        L.ppi_adjoint_create.argtypes = [P(C.c_void_p), C.c_int32]
        L.ppi_adjoint_destroy.argtypes = [C.c_void_p]
        L.ppi_adjoint_destroy.restype = None
        L.ppi_adjoint_configure.argtypes = [C.c_void_p, P(Config), P(Symbol), C.c_int32]
        L.ppi_adjoint_set_dataset_rowmajor.argtypes = [C.c_void_p, P(C.c_float), C.c_int32, C.c_int32]
        L.ppi_adjoint_forward.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_float), P(C.c_int32), C.c_int32,
                                          C.c_int32, C.c_float, P(C.c_float)]
        L.ppi_adjoint_backward.argtypes = [C.c_void_p, P(C.c_float), C.c_int32, P(C.c_float)]
        L.ppi_adjoint_descend.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_int32, P(C.c_float), P(C.c_float)]
        L.ppi_adjoint_time.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_double)]
        L.ppi_adjoint_get_stats.argtypes = [C.c_void_p, P(AdjointStats)]
        L.ppi_adjoint_set_rows_per_lane.argtypes = [C.c_void_p, C.c_int32]
        L.ppi_ef_evaluate.argtypes = [C.c_char_p, P(C.c_float), P(C.c_float), C.c_int32, P(C.c_float)]
        for name in ENGINE_SYMBOLS:
            if name not in ("ppi_abi_version", "ppi_last_error", "ppi_engine_destroy", "ppi_batch_destroy",
                            "ppi_adjoint_destroy"):
                getattr(L, name).restype = C.c_int32
        _lib = L
    return _lib


def wl() -> C.CDLL:
    global _wl
    if _wl is None:
        if not os.path.exists(WORKLOAD_LIB):
            raise RuntimeError(f"{WORKLOAD_LIB} is missing: run `make -C {HERE}`")
        L = C.CDLL(WORKLOAD_LIB)
        P = C.POINTER
        L.ppi_wl_grow.argtypes = [P(WlGrammar), C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  P(C.c_int32), P(C.c_float), P(C.c_int32)]
        L.ppi_wl_grow.restype = C.c_int64
        L.ppi_wl_exact.argtypes = [P(WlGrammar), C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   P(C.c_int32), P(C.c_float), P(C.c_int32)]
        L.ppi_wl_exact.restype = C.c_int64
        L.ppi_wl_dataset.argtypes = [C.c_int32, C.c_uint64, C.c_int32, C.c_int32, P(C.c_float)]
        L.ppi_wl_dataset.restype = C.c_int32
        _wl = L
    return _wl


def device_count() -> int:
    n = C.c_int32(0)
    st = lib().ppi_device_count(C.byref(n))
    return n.value if st == OK else 0


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _check(st: int) -> None:
    if st != OK:
        raise PpiError(st, lib().ppi_last_error().decode(errors="replace"))


class Engine:
    """One engine on one HIP device (acc_interpret_init + acc_interpret equivalent)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().ppi_engine_create(C.byref(self._h), device))
        self._cost = None
        self._children = []  # objects that evaluate through this engine (GE): closed before it

    def close(self) -> None:
        for ref in self._children:
            child = ref()
            if child is not None:
                child.close()
        self._children = []
        if self._h:
            lib().ppi_engine_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def configure(self, table: Sequence[Tuple[int, int, float]], strategy: int = SEQ, error_fn: int = EF_ABS,
                  reducemax: bool = False, protected: bool = True, native: bool = False,
                  cost: Optional[np.ndarray] = None, form: int = FORM_AUTO, max_local_size: int = 0,
                  compute_units: int = 0, rows_per_lane: int = 0, programs_per_tile: int = 0,
                  ordered_pass_rows: int = 0, fold_constants: bool = True, ef_expr: Optional[str] = None) -> None:
        """ef_expr: the text of a user ERROR(X,Y) (error_fn becomes EF_EXPR)."""
        cfg = Config()
        if ef_expr is not None:
            error_fn = EF_EXPR
            self._ef_text = ef_expr.encode()
            cfg.ef_expr = self._ef_text
        cfg.strategy, cfg.form, cfg.error_fn = strategy, form, error_fn
        cfg.reducemax, cfg.protected_ops, cfg.native_math = int(reducemax), int(protected), int(native)
        if cost is not None:
            self._cost = np.ascontiguousarray(cost, dtype=np.float32)
            cfg.cost_n = self._cost.shape[0]
            cfg.cost = _ptr(self._cost, C.c_float)
        cfg.max_local_size, cfg.compute_units = max_local_size, compute_units
        cfg.rows_per_lane, cfg.programs_per_tile = rows_per_lane, programs_per_tile
        cfg.ordered_pass_rows = ordered_pass_rows
        cfg.no_constant_folding = 0 if fold_constants else 1
        tab = (Symbol * max(1, len(table)))()
        for k, (s, op, v) in enumerate(table):
            tab[k].symbol, tab[k].op, tab[k].value = s, op, v
        _check(lib().ppi_engine_configure(self._h, C.byref(cfg), tab, len(table)))

    def set_dataset(self, data: np.ndarray) -> None:
        """Row-major nlin x ncol float32 dataset; target in the last column."""
        d = np.ascontiguousarray(data, dtype=np.float32)
        _check(lib().ppi_engine_set_dataset_rowmajor(self._h, _ptr(d, C.c_float), d.shape[0], d.shape[1]))

    def set_dataset_rows(self, data: np.ndarray) -> None:
        """Same data passed as float** (one pointer per row), the reference's layout."""
        d = np.ascontiguousarray(data, dtype=np.float32)
        rows = (C.POINTER(C.c_float) * d.shape[0])()
        base = d.ctypes.data
        for r in range(d.shape[0]):
            rows[r] = C.cast(base + r * d.shape[1] * 4, C.POINTER(C.c_float))
        _check(lib().ppi_engine_set_dataset(self._h, rows, d.shape[0], d.shape[1]))

    def upload(self, pop: "Population") -> None:
        self._nprog = pop.n
        _check(lib().ppi_engine_upload(self._h, _ptr(pop.phenotype, C.c_int32), _ptr(pop.ephemeral, C.c_float),
                                       _ptr(pop.size, C.c_int32), pop.stride, pop.n))

    def launch(self) -> None:
        _check(lib().ppi_engine_launch(self._h))

    def wait(self, n: int, alpha: float = 0.0, best_size: int = 0):
        vec = np.empty(n, dtype=np.float32)
        idx = np.zeros(max(1, best_size), dtype=np.int32)
        bs = C.c_int32(best_size)
        _check(lib().ppi_engine_wait(self._h, alpha, _ptr(vec, C.c_float),
                                     _ptr(idx, C.c_int32) if best_size else None,
                                     C.byref(bs) if best_size else None))
        return vec, idx[:bs.value] if best_size else None

    def evaluate(self, pop: "Population", alpha: float = 0.0, best_size: int = 0):
        self.upload(pop)
        self.launch()
        return self.wait(pop.n, alpha, best_size)

    def evaluate_noalloc(self, pop: "Population", alpha: float = 0.0, best_size: int = 0):
        """ppi_engine_evaluate_noalloc: the evaluation with every device buffer in the library's static
        scratch array (no device allocation); the resident program set is left as it was."""
        vec = np.empty(pop.n, dtype=np.float32)
        idx = np.zeros(max(1, best_size), dtype=np.int32)
        bs = C.c_int32(best_size)
        _check(lib().ppi_engine_evaluate_noalloc(self._h, _ptr(pop.phenotype, C.c_int32),
                                                 _ptr(pop.ephemeral, C.c_float), _ptr(pop.size, C.c_int32),
                                                 pop.stride, pop.n, alpha, _ptr(vec, C.c_float),
                                                 _ptr(idx, C.c_int32) if best_size else None,
                                                 C.byref(bs) if best_size else None))
        return vec, idx[:bs.value] if best_size else None

    def evaluate_ranked(self, pop: "Population", alpha: float = 0.0, best_size: int = 0):
        """ppi_engine_evaluate_ranked: alpha*size and the strategy's ranking computed on the device."""
        vec = np.empty(pop.n, dtype=np.float32)
        idx = np.zeros(max(1, best_size), dtype=np.int32)
        bs = C.c_int32(best_size)
        _check(lib().ppi_engine_evaluate_ranked(self._h, _ptr(pop.phenotype, C.c_int32),
                                                _ptr(pop.ephemeral, C.c_float), _ptr(pop.size, C.c_int32),
                                                pop.stride, pop.n, alpha, _ptr(vec, C.c_float),
                                                _ptr(idx, C.c_int32) if best_size else None,
                                                C.byref(bs) if best_size else None))
        return vec, idx[:bs.value] if best_size else None

    def resident_tape(self):
        """(meta n x 4 uint32, code m x 2 uint32) of the resident compiled programs."""
        n = C.c_int64(0)
        _check(lib().ppi_engine_resident_tape(self._h, None, None, C.byref(n)))
        nprog = self._nprog
        meta = np.zeros((nprog, 4), np.uint32)
        code = np.zeros((n.value, 2), np.uint32)
        _check(lib().ppi_engine_resident_tape(self._h, _ptr(meta, C.c_uint32), _ptr(code, C.c_uint32), C.byref(n)))
        return meta, code

    def predict(self, phenotype: np.ndarray, ephemeral: np.ndarray, size: int, nlin: int) -> np.ndarray:
        ph = np.ascontiguousarray(phenotype, dtype=np.int32)
        ep = np.ascontiguousarray(ephemeral, dtype=np.float32)
        out = np.empty(nlin, dtype=np.float32)
        _check(lib().ppi_engine_predict(self._h, _ptr(ph, C.c_int32), _ptr(ep, C.c_float), size,
                                        _ptr(out, C.c_float)))
        return out

    def stats(self) -> Stats:
        s = Stats()
        _check(lib().ppi_engine_get_stats(self._h, C.byref(s)))
        return s

    def time(self, iters: int) -> float:
        ms = C.c_double(0.0)
        _check(lib().ppi_engine_time(self._h, iters, C.byref(ms)))
        return ms.value


def make_config(strategy=SEQ, error_fn=EF_ABS, reducemax=False, protected=True, native=False, cost=None,
                form=FORM_AUTO, max_local_size=0, compute_units=0, rows_per_lane=0, programs_per_tile=0,
                fold_constants=True, ef_expr=None):
    """ef_expr: the text of a user ERROR(X,Y) (error_fn becomes EF_EXPR)."""
    cfg = Config()
    if ef_expr is not None:
        error_fn = EF_EXPR
    cfg.strategy, cfg.form, cfg.error_fn = strategy, form, error_fn
    cfg.reducemax, cfg.protected_ops, cfg.native_math = int(reducemax), int(protected), int(native)
    keep = None
    if cost is not None:
        keep = np.ascontiguousarray(cost, dtype=np.float32)
        cfg.cost_n = keep.shape[0]
        cfg.cost = _ptr(keep, C.c_float)
    cfg.max_local_size, cfg.compute_units = max_local_size, compute_units
    cfg.rows_per_lane, cfg.programs_per_tile = rows_per_lane, programs_per_tile
    cfg.no_constant_folding = 0 if fold_constants else 1
    if ef_expr is not None:
        text = ef_expr.encode()
        cfg.ef_expr = text
        keep = (keep, text)
    return cfg, keep


def ef_evaluate(text: str, X: np.ndarray, Y: np.ndarray) -> np.ndarray:
    """ppi_ef_evaluate: a user ERROR(X,Y) compiled and evaluated on the host, as the device does."""
    x = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(Y, dtype=np.float32)
    out = np.empty(x.shape[0], dtype=np.float32)
    _check(lib().ppi_ef_evaluate(text.encode(), _ptr(x, C.c_float), _ptr(y, C.c_float), x.shape[0],
                                 _ptr(out, C.c_float)))
    return out


def _table_struct(table):
    tab = (Symbol * max(1, len(table)))()
    for k, (s, op, v) in enumerate(table):
        tab[k].symbol, tab[k].op, tab[k].value = s, op, v
    return tab


class Batch:
    """Many (population, dataset) pairs evaluated in one persistent, graph-captured launch."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().ppi_batch_create(C.byref(self._h), device))
        self._keep = []
        self.sizes = []

    def close(self):
        if self._h:
            lib().ppi_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, table, **kw):
        cfg, keep = make_config(**kw)
        self._keep.append(keep)
        _check(lib().ppi_batch_configure(self._h, C.byref(cfg), _table_struct(table), len(table)))
        self.sizes = []

    def add_dataset(self, data: np.ndarray) -> int:
        d = np.ascontiguousarray(data, dtype=np.float32)
        i = C.c_int32(-1)
        _check(lib().ppi_batch_add_dataset(self._h, _ptr(d, C.c_float), d.shape[0], d.shape[1], C.byref(i)))
        return i.value

    def add_population(self, dataset_id: int, pop: "Population") -> int:
        i = C.c_int32(-1)
        _check(lib().ppi_batch_add_population(self._h, dataset_id, _ptr(pop.phenotype, C.c_int32),
                                              _ptr(pop.ephemeral, C.c_float), _ptr(pop.size, C.c_int32),
                                              pop.stride, pop.n, C.byref(i)))
        self.sizes.append(pop.n)
        return i.value

    def clear_populations(self):
        _check(lib().ppi_batch_clear_populations(self._h))
        self.sizes = []

    def set_direct_threshold(self, min_items: int):
        """Pairs with at least min_items (row tile, program) items run as their own FP launch at the
        next build(): 0 = default (the pair fills the device), negative = none, 1 = every pair."""
        _check(lib().ppi_batch_set_direct_threshold(self._h, min_items))

    def build(self):
        _check(lib().ppi_batch_build(self._h))

    def run(self):
        _check(lib().ppi_batch_launch(self._h))
        _check(lib().ppi_batch_wait(self._h))

    def results(self, pop_id: int, alpha: float = 0.0, best_size: int = 0):
        n = self.sizes[pop_id]
        vec = np.empty(n, dtype=np.float32)
        idx = np.zeros(max(1, best_size), dtype=np.int32)
        bs = C.c_int32(best_size)
        _check(lib().ppi_batch_results(self._h, pop_id, alpha, _ptr(vec, C.c_float),
                                       _ptr(idx, C.c_int32) if best_size else None,
                                       C.byref(bs) if best_size else None))
        return vec, (idx[:bs.value] if best_size else None)

    def time(self, iters: int) -> float:
        ms = C.c_double(0.0)
        _check(lib().ppi_batch_time(self._h, iters, C.byref(ms)))
        return ms.value

    def stats(self) -> BatchStats:
        s = BatchStats()
        _check(lib().ppi_batch_get_stats(self._h, C.byref(s)))
        return s

    def pair_shape(self, pop_id: int) -> Tuple[int, int, int]:
        """(rows per lane, LDS feature slots, register-stack slots) the autotuner gave a population."""
        r, f, s = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        _check(lib().ppi_batch_pair_shape(self._h, pop_id, C.byref(r), C.byref(f), C.byref(s)))
        return r.value, f.value, s.value


class Adjoint:
    """d(fitness)/d(ephemeral constants): forward, backward and on-device descent."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().ppi_adjoint_create(C.byref(self._h), device))
        self._keep = []
        self.pop = None

    def close(self):
        if self._h:
            lib().ppi_adjoint_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, table, **kw):
        cfg, keep = make_config(**kw)
        self._keep.append(keep)
        _check(lib().ppi_adjoint_configure(self._h, C.byref(cfg), _table_struct(table), len(table)))

    def set_dataset(self, data: np.ndarray) -> None:
        d = np.ascontiguousarray(data, dtype=np.float32)
        _check(lib().ppi_adjoint_set_dataset_rowmajor(self._h, _ptr(d, C.c_float), d.shape[0], d.shape[1]))

    def forward(self, pop: "Population", alpha: float = 0.0) -> np.ndarray:
        fit = np.empty(pop.n, dtype=np.float32)
        _check(lib().ppi_adjoint_forward(self._h, _ptr(pop.phenotype, C.c_int32), _ptr(pop.ephemeral, C.c_float),
                                         _ptr(pop.size, C.c_int32), pop.stride, pop.n, alpha, _ptr(fit, C.c_float)))
        self.pop = pop
        return fit

    def backward(self, grad_fitness: Optional[np.ndarray] = None, partials: int = PARTIALS_AUTO) -> np.ndarray:
        g = np.empty(self.pop.n * self.pop.stride, dtype=np.float32)
        gf = None if grad_fitness is None else np.ascontiguousarray(grad_fitness, dtype=np.float32)
        _check(lib().ppi_adjoint_backward(self._h, _ptr(gf, C.c_float) if gf is not None else None, partials,
                                          _ptr(g, C.c_float)))
        return g

    def descend(self, steps: int, step: float = 0.1, partials: int = PARTIALS_AUTO):
        eph = np.empty(self.pop.n * self.pop.stride, dtype=np.float32)
        fit = np.empty(self.pop.n, dtype=np.float32)
        _check(lib().ppi_adjoint_descend(self._h, steps, step, partials, _ptr(eph, C.c_float), _ptr(fit, C.c_float)))
        return eph, fit

    def time(self, iters: int, partials: int = PARTIALS_AUTO) -> float:
        ms = C.c_double(0.0)
        _check(lib().ppi_adjoint_time(self._h, iters, partials, C.byref(ms)))
        return ms.value

    def stats(self) -> AdjointStats:
        s = AdjointStats()
        _check(lib().ppi_adjoint_get_stats(self._h, C.byref(s)))
        return s

    def set_rows_per_lane(self, rows: int) -> None:
        """0: chosen by measurement (default), 1, or 2 (programs of at most 5 stack and adjoint slots)."""
        _check(lib().ppi_adjoint_set_rows_per_lane(self._h, rows))


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

class WlGrammar(C.Structure):
    _fields_ = [
        ("binary_ops", C.POINTER(C.c_int32)), ("n_binary", C.c_int32),
        ("unary_ops", C.POINTER(C.c_int32)), ("n_unary", C.c_int32),
        ("t_attribute", C.c_int32), ("t_const", C.c_int32), ("n_attributes", C.c_int32),
        ("const_lo", C.c_float), ("const_hi", C.c_float), ("p_const", C.c_float), ("p_leaf", C.c_float),
        ("max_consts", C.c_int32),
    ]


@dataclasses.dataclass
class Population:
    phenotype: np.ndarray  # int32 [n*stride]
    ephemeral: np.ndarray  # float32 [n*stride]
    size: np.ndarray       # int32 [n]
    stride: int

    @property
    def n(self) -> int:
        return int(self.size.shape[0])

    @property
    def total_size(self) -> int:
        return int(self.size.astype(np.int64).sum())

    def subset(self, k: int) -> "Population":
        return Population(self.phenotype[: k * self.stride].copy(), self.ephemeral[: k * self.stride].copy(),
                          self.size[:k].copy(), self.stride)


BINARY = ["+", "-", "*", "/"]
UNARY = ["sin", "cos", "exp", "log", "sqrt"]


def config_bnf(n_attributes: int, binary=BINARY, unary=UNARY, const: bool = True) -> str:
    """The grammar of the benchmark populations, in PPI's BNF format."""
    terms = " | ".join([f"attr{k}" for k in range(n_attributes)] + (["const"] if const else []))
    return ("S = <expr>\n"
            "P = <expr> ::= <bin> <expr> <expr> | <un> <expr> | <term>\n"
            f"    <bin> ::= {' | '.join(binary)}\n"
            f"    <un> ::= {' | '.join(unary)}\n"
            f"    <term> ::= {terms}\n")


class Grammar:
    """Symbol numbering + engine table for a benchmark grammar."""

    def __init__(self, n_attributes: int, binary=BINARY, unary=UNARY, const: bool = True):
        self.bnf = config_bnf(n_attributes, binary, unary, const)
        self.g = _grammar.parse_bnf(self.bnf)
        self.values = self.g.values()
        self.table = self.g.engine_table()
        self.n_attributes = n_attributes
        self.binary = [self.values[_grammar._terminal_name(t)] for t in binary]
        self.unary = [self.values[_grammar._terminal_name(t)] for t in unary]
        self.t_attribute = self.values["T_ATTRIBUTE"]
        self.t_const = self.values.get("T_CONST", -1)

    def wl(self, const_lo=-5.0, const_hi=5.0, p_const=0.3, p_leaf=0.3, max_consts=0):
        g = WlGrammar()
        self._b = np.array(self.binary, dtype=np.int32)
        self._u = np.array(self.unary, dtype=np.int32)
        g.binary_ops, g.n_binary = _ptr(self._b, C.c_int32), len(self.binary)
        g.unary_ops, g.n_unary = _ptr(self._u, C.c_int32), len(self.unary)
        g.t_attribute, g.t_const, g.n_attributes = self.t_attribute, self.t_const, self.n_attributes
        g.const_lo, g.const_hi, g.p_const, g.p_leaf, g.max_consts = const_lo, const_hi, p_const, p_leaf, max_consts
        return g


def grow(gr: Grammar, n: int, seed: int, max_depth: int, max_symbols: int, stride: Optional[int] = None,
         **kw) -> Population:
    stride = stride or max_symbols
    ph = np.zeros(n * stride, dtype=np.int32)
    ep = np.zeros(n * stride, dtype=np.float32)
    sz = np.zeros(n, dtype=np.int32)
    g = gr.wl(**kw)
    r = wl().ppi_wl_grow(C.byref(g), seed, n, stride, max_depth, max_symbols, _ptr(ph, C.c_int32),
                         _ptr(ep, C.c_float), _ptr(sz, C.c_int32))
    if r < 0:
        raise ValueError("ppi_wl_grow failed")
    return Population(ph, ep, sz, stride)


def exact(gr: Grammar, n: int, seed: int, len_lo: int, len_hi: int, stride: Optional[int] = None, **kw) -> Population:
    stride = stride or len_hi
    ph = np.zeros(n * stride, dtype=np.int32)
    ep = np.zeros(n * stride, dtype=np.float32)
    sz = np.zeros(n, dtype=np.int32)
    g = gr.wl(**kw)
    r = wl().ppi_wl_exact(C.byref(g), seed, n, stride, len_lo, len_hi, _ptr(ph, C.c_int32),
                          _ptr(ep, C.c_float), _ptr(sz, C.c_int32))
    if r < 0:
        raise ValueError("ppi_wl_exact failed")
    return Population(ph, ep, sz, stride)


DS_FORMULA_SMALL, DS_POLY_NORMAL, DS_CLUSTERS, DS_POLY_UNIFORM, DS_FORMULA_ADJOINT = 1, 2, 3, 4, 5


def dataset(kind: int, seed: int, nlin: int, nfeat: int) -> np.ndarray:
    out = np.zeros((nlin, nfeat + 1), dtype=np.float32)
    if wl().ppi_wl_dataset(kind, seed, nlin, nfeat, _ptr(out, C.c_float)) != 0:
        raise ValueError("ppi_wl_dataset failed")
    return out


COST4 = (1.0 - np.eye(4)).astype(np.float32)


@dataclasses.dataclass
class Workload:
    name: str
    grammar: Grammar
    pop: Population
    data: np.ndarray
    error_fn: int
    cost: Optional[np.ndarray] = None
    reducemax: bool = False
    protected: bool = True

    @property
    def node_evals(self) -> float:
        return float(self.pop.total_size) * self.data.shape[0]


def workload(name: str, n_programs: Optional[int] = None, nlin: Optional[int] = None, seed: int = 1) -> Workload:
    """Seeded workloads of the benchmark configurations (sizes can be scaled down for tests).

    small            1,000 programs, depth <= 6, <= 64 symbols, 1,000 x 4 U[-1,1],
                     y = x0^2 - 0.5*x1*x2 + sin(x3), EF = |X-Y|
    regression       30,000 programs, depth <= 10, <= 128 symbols, 16 attributes,
                     100,000 x 16 N(0,1), y = random cubic + noise, EF = (X-Y)^2
    classification   30,000 programs (depth <= 10, <= 128), 150,000 x 4 from 4 clusters,
                     EF = COST[(int)clamp(X,0,3)][(int)Y] with a 0/1 cost matrix
    adjoint          30,000 programs (depth <= 10, <= 128 symbols, <= 8 T_CONST each, constants
                     U[-5,5]), 1,000,000 x 8 N(0,1), y = exp(0.5*x0) + sin(2*x1) + x2*x3*x4,
                     EF = (X-Y)^2
    """
    if name == "small":
        gr = Grammar(4)
        pop = grow(gr, n_programs or 1000, seed, 6, 64)
        data = dataset(DS_FORMULA_SMALL, seed + 1, nlin or 1000, 4)
        return Workload(name, gr, pop, data, EF_ABS)
    if name == "regression":
        gr = Grammar(16)
        pop = grow(gr, n_programs or 30000, seed, 10, 128)
        data = dataset(DS_POLY_NORMAL, seed + 1, nlin or 100000, 16)
        return Workload(name, gr, pop, data, EF_SQUARED)
    if name in ("classification", "classification_max"):
        gr = Grammar(4)
        pop = grow(gr, n_programs or 30000, seed, 10, 128)
        data = dataset(DS_CLUSTERS, seed + 1, nlin or 150000, 4)
        return Workload(name, gr, pop, data, EF_COST_CLAMP, cost=COST4, reducemax=name.endswith("_max"))
    if name == "adjoint":
        gr = Grammar(8)
        pop = grow(gr, n_programs or 30000, seed, 10, 128, max_consts=8)
        data = dataset(DS_FORMULA_ADJOINT, seed + 1, nlin or 1000000, 8)
        return Workload(name, gr, pop, data, EF_SQUARED)
    raise ValueError(f"unknown workload {name!r}")


@dataclasses.dataclass
class BatchedWorkload:
    grammars: List[Grammar]
    pops: List[Population]
    datas: List[np.ndarray]
    error_fn: int = EF_SQUARED

    @property
    def node_evals(self) -> float:
        return float(sum(p.total_size * d.shape[0] for p, d in zip(self.pops, self.datas)))


def batched_workload(n_pairs: int = 64, seed: int = 4, p_range=(500, 20000), len_range=(8, 256),
                     rows_range=(1000, 1000000), feat_range=(2, 32)) -> BatchedWorkload:
    """Uneven (population, dataset) pairs: population size and row count log-uniform,
    exact program lengths uniform, 2..32 U[-1,1] features with a random polynomial target."""
    rng = np.random.default_rng(seed)
    grammars, pops, datas = [], [], []
    for k in range(n_pairs):
        P = int(round(math.exp(rng.uniform(math.log(p_range[0]), math.log(p_range[1])))))
        nlin = int(round(math.exp(rng.uniform(math.log(rows_range[0]), math.log(rows_range[1])))))
        nfeat = int(rng.integers(feat_range[0], feat_range[1] + 1))
        gr = Grammar(nfeat)
        pops.append(exact(gr, P, seed * 1000 + 2 * k, len_range[0], len_range[1]))
        datas.append(dataset(DS_POLY_UNIFORM, seed * 1000 + 2 * k + 1, nlin, nfeat))
        grammars.append(gr)
    return BatchedWorkload(grammars, pops, datas)


class GE:
    """Grammatical evolution on the device (ppi_ge_*), evaluating through an Engine."""

    def __init__(self, engine: Engine, setup: "GeSetup"):
        import weakref
        self._h = C.c_void_p()
        self.engine = engine
        self.setup = setup
        _check(lib().ppi_ge_create(C.byref(self._h), engine._h))
        engine._children.append(weakref.ref(self))
        _check(lib().ppi_ge_configure(self._h, C.byref(setup.params), C.byref(setup.grammar)))
        self.P = setup.params.population_size
        self.nb = setup.params.number_of_bits

    def close(self):
        if self._h:
            lib().ppi_ge_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self):
        _check(lib().ppi_ge_initialize(self._h))

    def evolve(self, generations: int = 1):
        _check(lib().ppi_ge_evolve(self._h, generations))

    def population(self):
        g = np.zeros((self.P, self.nb), np.uint8)
        f = np.zeros(self.P, np.float32)
        _check(lib().ppi_ge_get_population(self._h, _ptr(g, C.c_uint8), _ptr(f, C.c_float)))
        return g, f

    def set_immigrants(self, genomes: np.ndarray):
        g = np.ascontiguousarray(genomes, dtype=np.uint8)
        _check(lib().ppi_ge_set_immigrants(self._h, _ptr(g, C.c_uint8), g.shape[0]))

    def best(self):
        f = C.c_float(0.0)
        g = np.zeros(self.nb, np.uint8)
        _check(lib().ppi_ge_best(self._h, C.byref(f), _ptr(g, C.c_uint8)))
        return f.value, g

    def decoded(self):
        st = self.stats()
        mx = st.max_size_phenotype
        ph = np.zeros((self.P, mx), np.int32)
        ep = np.zeros((self.P, mx), np.float32)
        sz = np.zeros(self.P, np.int32)
        stride = C.c_int32(0)
        _check(lib().ppi_ge_decoded(self._h, _ptr(ph, C.c_int32), _ptr(ep, C.c_float), _ptr(sz, C.c_int32),
                                    C.byref(stride)))
        return ph, ep, sz

    def tape(self):
        n = C.c_int64(0)
        _check(lib().ppi_ge_tape(self._h, None, None, C.byref(n)))
        meta = np.zeros((self.P, 4), np.uint32)
        code = np.zeros((n.value, 2), np.uint32)
        _check(lib().ppi_ge_tape(self._h, _ptr(meta, C.c_uint32), _ptr(code, C.c_uint32), C.byref(n)))
        return meta, code

    def stats(self) -> GeStats:
        s = GeStats()
        _check(lib().ppi_ge_get_stats(self._h, C.byref(s)))
        return s
