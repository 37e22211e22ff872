"""Exhaustive check of the engine's correctly rounding float exp, log, sin and cos.

Each routine (ppi-auto_amd/csrc/exact_math.h) is compared with the device library's
double routine rounded to float — the engine's previous implementation of the
sequential path's (float) f((double) x) (functions.h, interpreter_core:103-117,
140-142) — and sqrt with the library's correctly rounded sqrtf, on every one of the
2^32 float bit patterns, NaNs compared as NaN.
Identical results on all inputs make the two implementations interchangeable;
their agreement with the oracle (glibc's double routines, rounded) is covered by
the parity tests.
"""
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "devmath", "libppi_mathcheck.so")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mathcheck():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB)
    lib.ppi_mathcheck.argtypes = [C.c_int, C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
    lib.ppi_mathcheck.restype = C.c_int
    return lib


@pytest.mark.parametrize("fn,name", [(0, "exp"), (1, "log"), (2, "sin"), (3, "cos"), (4, "sqrt")])
def test_all_float_inputs(mathcheck, fn, name):
    bad = C.c_ulonglong(0)
    samples = (C.c_uint32 * 64)()
    ns = C.c_uint32(0)
    ms = C.c_float(0)
    print(f"exhaustive {name}: 2^32 inputs", flush=True)
    rc = mathcheck.ppi_mathcheck(fn, C.byref(bad), samples, C.byref(ns), C.byref(ms))
    assert rc == 0, f"HIP error {rc}"
    xs = np.array(samples[: min(64, ns.value)], dtype=np.uint32).view(np.float32)
    print(f"{name}: {bad.value} mismatches in {ms.value:.1f} ms; first inputs {xs[:8]}", flush=True)
    assert bad.value == 0, f"{name}: {bad.value} inputs differ, e.g. {[float(v) for v in xs[:8]]}"
