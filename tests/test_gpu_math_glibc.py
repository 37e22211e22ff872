"""Device exp, log, sin, cos, tan, exp10, exp2, log10, log2 and sqrt against the sequential path's
glibc results, bit for bit.

The sequential path evaluates these operators as (float) f((double) x) with glibc's
double routines (sequential.cc:117-209 with functions.h:131-147 and
interpreter_core:103-117,140-142); the oracle does the same.  Each operator is run
on one column of a million inputs in prediction mode (ppp.cc:98-151): random bit
patterns over every float exponent and sign, typical magnitudes, and the special
values (signed zeros, infinities, NaN, denormals, the largest finite floats,
multiples of pi/2).  Both the plain operator (which the engine may serve from a
derived column) and the operator over a computed value (x0 * 1, evaluated inside
the interpreter) are checked, with PROTECTED on and off.

A second set runs each transcendental on every float input whose glibc double result lies within
256 double ulps of a float rounding boundary (tests/golden/math_hard_inputs.npz, made by
tests/golden/make_math_hard_inputs.py from a scan of all 2^32 inputs): the inputs where a double
routine a few ulps from glibc's could round to another float, and where the exact float routines
leave their fast path for the device library's double routine.
"""
import os

import numpy as np
import pytest

import oracle
import ppiamd
from helpers import Builder

pytestmark = pytest.mark.gpu


def inputs(seed, n_random=600_000, n_typical=400_000):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 1 << 32, n_random, dtype=np.uint64).astype(np.uint32)
    typical = (rng.standard_normal(n_typical) * np.exp(rng.uniform(-6, 6, n_typical))).astype(np.float32)
    half_pi = np.float32(np.pi / 2)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 1.17549435e-38, 3.4028235e38,
                        -3.4028235e38, 1.0, -1.0, 88.72283, 88.72284, -87.33655, -103.97208, -104.0, 128.0,
                        16777216.0, 16777215.0, 505.796417, half_pi, -half_pi, 3 * half_pi, np.float32(np.pi)],
                       dtype=np.float32)
    return np.concatenate([bits.view(np.float32), typical, special])


@pytest.fixture(scope="module")
def eng():
    if ppiamd.device_count() == 0:
        pytest.fail("no HIP device visible")
    e = ppiamd.Engine(0)
    yield e
    e.close()


OPS = ["exp", "log", "sin", "cos", "sqrt", "tan", "exp10", "exp2", "log10", "log2"]
HARD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "math_hard_inputs.npz"))


def compare(eng, op, prot, x, x1=None, progs=None):
    x1 = np.zeros_like(x) if x1 is None else x1
    data = np.stack([x, x1, np.zeros_like(x)], axis=1).astype(np.float32)
    b = Builder(2)
    print(f"{op} protected={prot}: {x.size} inputs", flush=True)
    eng.configure(b.table, strategy=ppiamd.SEQ, error_fn=ppiamd.EF_ABS, protected=prot, compute_units=256)
    eng.set_dataset(data)
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS, protected=prot)
    for prog in progs or ([op, "x0"], [op, "*", "x0", 1.0]):
        ph, ep = b.encode(prog)
        out = eng.predict(np.array(ph, np.int32), np.array(ep, np.float32), len(ph), data.shape[0])
        ref, _ = oracle.interpret(cfg, b.table, b.population([prog]), data, ppp_mode=1, prediction_mode=1)
        same = (out.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(out) & np.isnan(ref))
        bad = np.nonzero(~same)[0]
        assert bad.size == 0, (f"{prog}: {bad.size} of {x.size} differ, e.g. x={x[bad[:5]]} "
                               f"device={out[bad[:5]]} glibc={ref[bad[:5]]}")


@pytest.mark.parametrize("prot", [True, False])
@pytest.mark.parametrize("op", OPS)
def test_unary_matches_glibc(eng, op, prot):
    compare(eng, op, prot, inputs(seed=2 * OPS.index(op) + int(prot)))


@pytest.mark.parametrize("prot", [True, False])
@pytest.mark.parametrize("op", [op for op in OPS if op != "sqrt"])
def test_hard_inputs_match_glibc(eng, op, prot):
    x = HARD[op].view(np.float32)
    assert x.size > 400
    compare(eng, op, prot, x)


@pytest.mark.parametrize("prot", [True, False])
@pytest.mark.parametrize("op", ["pow", "mod", "gamma"])
def test_pow_mod_gamma_match_glibc(eng, op, prot):
    """pow and fmod of two inputs and GAMMA (which calls sinh and pow) over the sequential path's
    double routines (interpreter_core:68-75,127-129)."""
    rng = np.random.default_rng(100 + 2 * ["pow", "mod", "gamma"].index(op) + int(prot))
    x = inputs(seed=50 + int(prot), n_random=300_000, n_typical=200_000)
    if op == "gamma":
        x = np.concatenate([x, rng.uniform(-40, 40, 200_000).astype(np.float32)])
        compare(eng, op, prot, x)
        return
    n = x.size
    x1 = np.concatenate([(rng.standard_normal(n // 2) * np.exp(rng.uniform(-4, 4, n // 2))).astype(np.float32),
                         rng.integers(-40, 41, n - n // 2).astype(np.float32)])
    rng.shuffle(x1)
    if op == "pow":
        x = np.where(rng.random(n) < 0.5, np.abs(x), x).astype(np.float32)
    compare(eng, op, prot, x, x1, progs=[[op, "x0", "x1"], [op, "*", "x0", 1.0, "x1"]])
