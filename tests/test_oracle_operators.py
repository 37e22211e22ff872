"""Pins every operator of the CPU oracle (oracle/ppi_oracle.c) on its own.

The oracle restates interpreter_core:1-190 and functions.h:19-193.  Each case
below states what the reference's line fixes, independently of the oracle's
code: truth tables of the C `(bool)` cast, operand order (the first child in
prefix order is the left operand, interpreter_core:48-51), exact small-integer
results, IEEE special values (signed zeros, NaN, overflow), the values of the
double routines at exact points (exp2(10) = 1024, log10(1000) = 3, ...), the
protected points of functions.h:99-193, and the C library's own routines
called directly through ctypes where the reference calls them (fmax, fmin,
fmod, pow, round).  A transcription slip in one operator of the oracle — a
swapped operand, a missing protection, rint for round, a dropped promotion —
fails at least one case here.
"""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest

import oracle
import ppiamd
from helpers import Builder

F32 = np.float32
NAN = float("nan")
INF = float("inf")
FLT_MAX = float(np.finfo(np.float32).max)

NAMED = ["pi", "pi_2", "pi_4", "1_pi", "2_pi", "2_sqrtpi", "sqrt2", "sqrt1_2", "e", "log2e", "log10e", "ln2",
         "ln10"]
INTS = ["0", "1", "2", "7", "100", "n1", "n7", "n100"]

_B = Builder(4, extra_terminals=tuple(NAMED + INTS))
_LIBM = ctypes.CDLL(ctypes.util.find_library("m"))
for _f in ("fmax", "fmin", "fmod", "pow", "round", "sinh", "tan", "exp10", "exp2", "log10", "log2"):
    getattr(_LIBM, _f).restype = ctypes.c_double
    getattr(_LIBM, _f).argtypes = [ctypes.c_double] * (2 if _f in ("fmax", "fmin", "fmod", "pow") else 1)


def ev(tokens, protected=True, row=(0.0, 0.0, 0.0, 0.0, 0.0)):
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS, protected=protected)
    ph, ep = _B.encode(tokens)
    return F32(oracle.eval_row(cfg, _B.table, ph, ep, np.array(row, np.float32)))


def same(got, want):
    """Bitwise equality of float32 values (NaN == NaN, -0 != +0)."""
    g, w = np.array([got], np.float32), np.array([want], np.float32)
    if np.isnan(w[0]):
        return bool(np.isnan(g[0]))
    return g.view(np.uint32)[0] == w.view(np.uint32)[0]


def truth(x):
    """C's (bool)x for a float: true unless x compares equal to zero (NaN is true)."""
    return not (x == 0.0)


BOOL_VALUES = [-1.0, -0.0, 0.0, 0.5, NAN]


# --- logic (interpreter_core:5-19) -----------------------------------------------------------

@pytest.mark.parametrize("a", BOOL_VALUES)
@pytest.mark.parametrize("b", BOOL_VALUES)
def test_logic_truth_tables(a, b):
    ta, tb = truth(a), truth(b)
    assert same(ev(["and", a, b]), 1.0 if (ta and tb) else 0.0)
    assert same(ev(["or", a, b]), 1.0 if (ta or tb) else 0.0)
    assert same(ev(["xor", a, b]), 1.0 if (ta != tb) else 0.0)


@pytest.mark.parametrize("a", BOOL_VALUES)
def test_not(a):
    assert same(ev(["not", a]), 0.0 if truth(a) else 1.0)


def test_bool_of_values_is_not_truncation():
    # (bool)0.5 is true: a slip to (int)x would make it false
    assert same(ev(["and", 0.5, 0.25]), 1.0)
    assert same(ev(["not", 1e-30]), 0.0)


# --- relational (interpreter_core:20-43): left operand = first child ---------------------------

REL = {">": lambda a, b: a > b, ">=": lambda a, b: a >= b, "<": lambda a, b: a < b, "<=": lambda a, b: a <= b,
       "=": lambda a, b: a == b, "!=": lambda a, b: a != b}


@pytest.mark.parametrize("op", list(REL))
@pytest.mark.parametrize("a,b", [(2.0, 1.0), (1.0, 2.0), (1.0, 1.0), (-0.0, 0.0), (NAN, 1.0), (1.0, NAN),
                                 (NAN, NAN), (-INF, FLT_MAX)])
def test_relational(op, a, b):
    assert same(ev([op, a, b]), 1.0 if REL[op](a, b) else 0.0)


def test_relational_operand_order():
    assert same(ev([">", 3.0, 2.0]), 1.0) and same(ev([">", 2.0, 3.0]), 0.0)
    assert same(ev(["<", "x0", "x1"], row=(1, 2, 0, 0, 0)), 1.0)
    assert same(ev(["<", "x1", "x0"], row=(1, 2, 0, 0, 0)), 0.0)


# --- arithmetic (interpreter_core:44-75, 133-136) ---------------------------------------------

def test_sub_div_pow_mod_operand_order():
    assert same(ev(["-", 5.0, 3.0]), 2.0)
    assert same(ev(["/", 6.0, 3.0]), 2.0) and same(ev(["/", 3.0, 6.0]), 0.5)
    assert same(ev(["pow", 2.0, 3.0]), 8.0) and same(ev(["pow", 3.0, 2.0]), 9.0)
    assert same(ev(["mod", 5.0, 3.0]), 2.0) and same(ev(["mod", 3.0, 5.0]), 3.0)


def test_mean():
    assert same(ev(["mean", 1.0, 2.0]), 1.5)
    assert same(ev(["mean", -3.0, 3.0]), 0.0)
    # (a + b) / 2 in float: the sum overflows first (a/2 + b/2 would not)
    assert same(ev(["mean", FLT_MAX, FLT_MAX]), INF)
    # and the halving is exact: MEAN of two adjacent floats rounds the sum, not each half
    a = float(F32(1.0)); b = float(np.nextafter(F32(1.0), F32(2.0)))
    assert same(ev(["mean", a, b]), F32(F32(a + b) / F32(2.0)))


def test_protected_and_plain_division():
    assert same(ev(["/", 1.0, 0.0], protected=True), 1.0)
    assert same(ev(["/", 1.0, -0.0], protected=True), 1.0)  # -0 == 0
    assert same(ev(["/", 1.0, 0.0], protected=False), INF)
    assert same(ev(["/", -1.0, 0.0], protected=False), -INF)
    assert same(ev(["/", 0.0, 0.0], protected=False), NAN)
    assert same(ev(["/", 1.0, 3.0]), F32(1.0) / F32(3.0))


@pytest.mark.parametrize("a,b", [(1.0, 2.0), (2.0, 1.0), (NAN, 1.0), (1.0, NAN), (-0.0, 0.0), (0.0, -0.0),
                                 (-INF, 3.0), (NAN, NAN)])
def test_max_min_are_the_c_library_routines(a, b):
    """interpreter_core:60-67 call fmax/fmin of the C library on (promoted) floats."""
    assert same(ev(["max", a, b]), F32(_LIBM.fmax(a, b)))
    assert same(ev(["min", a, b]), F32(_LIBM.fmin(a, b)))


def test_max_min_nan_rule():
    # C99 F.9.9.2/3: a NaN operand is ignored
    assert same(ev(["max", NAN, -5.0]), -5.0) and same(ev(["min", 7.0, NAN]), 7.0)


@pytest.mark.parametrize("a,b,want", [(5.0, 3.0, 2.0), (-5.0, 3.0, -2.0), (5.0, -3.0, 2.0), (5.5, 2.0, 1.5),
                                      (5.0, 0.0, NAN), (INF, 2.0, NAN), (2.0, INF, 2.0), (-0.0, 3.0, -0.0)])
def test_mod_is_fmod(a, b, want):
    """fmod: sign of the dividend, exact remainder (C99 7.12.10.1)."""
    assert same(ev(["mod", a, b]), want)
    assert same(ev(["mod", a, b]), F32(_LIBM.fmod(a, b)))


@pytest.mark.parametrize("a,b,want", [(2.0, 10.0, 1024.0), (4.0, 0.5, 2.0), (-8.0, 1.0 / 3.0, NAN), (0.0, 0.0, 1.0),
                                      (NAN, 0.0, 1.0), (1.0, NAN, 1.0), (-2.0, 3.0, -8.0), (0.0, -1.0, INF),
                                      (10.0, 39.0, INF)])
def test_pow(a, b, want):
    """pow in double, rounded (C99 F.9.4.4 special cases; 10^39 overflows float)."""
    assert same(ev(["pow", a, b]), want)


def test_pow_is_double_rounded():
    # 3^0.3 via the double routine rounded to float, not powf
    for a, b in [(3.0, 0.3), (1.1, 123.5), (7.25, -2.75)]:
        fa, fb = float(F32(a)), float(F32(b))
        assert same(ev(["pow", fa, fb]), F32(_LIBM.pow(fa, fb)))


# --- unary (interpreter_core:76-151) -----------------------------------------------------------

def test_abs_neg():
    assert same(ev(["abs", -2.5]), 2.5) and same(ev(["abs", -0.0]), 0.0) and same(ev(["abs", NAN]), NAN)
    assert same(ev(["neg", 0.0]), -0.0) and same(ev(["neg", -3.0]), 3.0)


@pytest.mark.parametrize("op,x,want", [("pow2", 3.0, 9.0), ("pow3", -2.0, -8.0), ("pow4", 3.0, 81.0),
                                       ("pow5", 2.0, 32.0), ("pow5", -0.0, -0.0), ("pow2", -0.0, 0.0),
                                       ("pow2", 2.0 ** 64, INF), ("pow3", -(2.0 ** 43), -INF),
                                       ("pow4", 2.0 ** -40, 0.0), ("pow2", 1.5, 2.25)])
def test_integer_powers(op, x, want):
    """Repeated float multiplication (interpreter_core:79-90): exact on small integers, overflow to inf,
    underflow to 0 (2^-160 is below the smallest float denormal)."""
    assert same(ev([op, x]), want)


def test_pow4_pow5_are_left_to_right():
    """((x*x)*x)*x as written, each product rounded to float; x is chosen so that the grouping
    (x*x)*(x*x) rounds differently, so a regrouped restatement fails."""
    x = np.array([1066562514], np.uint32).view(np.float32)[0]  # 1.1441596
    chain4 = F32(F32(F32(x * x) * x) * x)
    pair4 = F32(F32(x * x) * F32(x * x))
    chain5 = F32(chain4 * x)
    pair5 = F32(pair4 * x)
    assert chain4 != pair4 and chain5 != pair5
    assert same(ev(["pow4", float(x)]), chain4)
    assert same(ev(["pow5", float(x)]), chain5)


@pytest.mark.parametrize("x,r,c,f", [(0.5, 1.0, 1.0, 0.0), (-0.5, -1.0, -0.0, -1.0), (1.5, 2.0, 2.0, 1.0),
                                     (2.5, 3.0, 3.0, 2.0), (-2.5, -3.0, -2.0, -3.0), (-0.0, -0.0, -0.0, -0.0),
                                     (1.2, 1.0, 2.0, 1.0), (-1.2, -1.0, -1.0, -2.0), (8388609.0, 8388609.0,
                                                                                      8388609.0, 8388609.0)])
def test_round_ceil_floor(x, r, c, f):
    """round: halves away from zero (C99 7.12.9.6; rint would give 2 for 2.5); ceil(-0.5) = -0."""
    assert same(ev(["round", x]), r)
    assert same(ev(["round", x]), F32(_LIBM.round(x)))
    assert same(ev(["ceil", x]), c)
    assert same(ev(["floor", x]), f)


@pytest.mark.parametrize("x,want", [(0.0, 1.0), (-0.0, 1.0), (1e-30, 1.0), (-1e-30, 0.0), (NAN, 0.0),
                                    (INF, 1.0), (-INF, 0.0)])
def test_step(x, want):
    """(x >= 0.0f) converted to float (interpreter_core:118-120)."""
    assert same(ev(["step", x]), want)


@pytest.mark.parametrize("x,want", [(3.0, 1.0), (-3.0, -1.0), (0.0, 0.0), (-0.0, 0.0), (NAN, 0.0),
                                    (1e-45, 1.0), (-INF, -1.0)])
def test_sign(x, want):
    assert same(ev(["sign", x]), want)


def test_logistic_and_gauss():
    assert same(ev(["logistic", 0.0]), 0.5)
    assert same(ev(["logistic", 100.0]), 1.0)   # 1 / (1 + e^-100): the sum rounds to 1
    assert same(ev(["logistic", -100.0]), 0.0)  # e^100 overflows float: 1 / inf
    assert same(ev(["logistic", 1.0]), F32(1.0) / F32(F32(1.0) + F32(math.exp(-1.0))))
    assert same(ev(["gauss", 0.0]), 1.0)
    assert same(ev(["gauss", 1.0]), F32(math.exp(-1.0)))
    assert same(ev(["gauss", -2.0]), F32(math.exp(-4.0)))
    assert same(ev(["gauss", 20.0]), 0.0)       # e^-400 underflows


@pytest.mark.parametrize("x,want", [(0.0, 1.0), (1.0, F32(math.e)), (-1.0, F32(math.exp(-1.0))), (89.0, INF),
                                    (-104.0, 0.0), (NAN, NAN)])
def test_exp(x, want):
    assert same(ev(["exp", x]), want)


@pytest.mark.parametrize("op,x,want", [("exp10", 2.0, 100.0), ("exp10", 0.0, 1.0), ("exp10", -1.0, F32(0.1)),
                                       ("exp10", 39.0, INF), ("exp2", 10.0, 1024.0), ("exp2", -1.0, 0.5),
                                       ("exp2", 0.5, F32(math.sqrt(2.0))), ("exp2", 128.0, INF),
                                       ("exp2", -149.0, 2.0 ** -149)])
def test_exp10_exp2(op, x, want):
    assert same(ev([op, x]), want)


@pytest.mark.parametrize("op,x,want", [("log", 1.0, 0.0), ("log", float(F32(math.e)), F32(math.log(float(F32(math.e))))),
                                       ("log10", 1000.0, 3.0), ("log10", 1.0, 0.0), ("log2", 8.0, 3.0),
                                       ("log2", 2.0 ** 100, 100.0), ("log", INF, INF)])
def test_logs_at_exact_points(op, x, want):
    assert same(ev([op, x], protected=True), want)
    assert same(ev([op, x], protected=False), want)


@pytest.mark.parametrize("op", ["log", "log10", "log2"])
def test_logs_protected_below_one(op):
    """functions.h:131-177: x < 1 -> 1 when PROTECTED, the double routine otherwise."""
    for x in [0.999, 0.5, 0.0, -0.0, -3.0]:
        assert same(ev([op, x], protected=True), 1.0)
    assert same(ev([op, NAN], protected=True), NAN)  # NaN < 1 is false
    assert same(ev(["log2", 0.25], protected=False), -2.0)
    assert same(ev(["log10", 0.01], protected=False), F32(_LIBM.log10(float(F32(0.01)))))
    assert same(ev(["log", 0.0], protected=False), -INF)
    assert same(ev(["log", -1.0], protected=False), NAN)


def test_sqrt():
    assert same(ev(["sqrt", 4.0]), 2.0) and same(ev(["sqrt", 2.0]), F32(math.sqrt(2.0)))
    assert same(ev(["sqrt", -1.0], protected=True), 1.0)
    assert same(ev(["sqrt", -0.0], protected=True), -0.0)  # -0 < 0 is false: sqrt(-0) = -0
    assert same(ev(["sqrt", -1.0], protected=False), NAN)
    assert same(ev(["sqrt", INF]), INF)


def test_sin_cos_special_values():
    assert same(ev(["sin", 0.0]), 0.0) and same(ev(["sin", -0.0]), -0.0)
    assert same(ev(["cos", 0.0]), 1.0) and same(ev(["cos", -0.0]), 1.0)
    half_pi = float(F32(math.pi / 2))
    assert same(ev(["sin", half_pi]), 1.0)
    assert same(ev(["cos", half_pi]), F32(math.cos(half_pi)))  # -4.37e-8, not 0
    assert same(ev(["sin", INF]), NAN) and same(ev(["cos", NAN]), NAN)


def test_tan():
    p2 = F32(1.57079637050629)  # M_PI_2_F
    for x in [float(p2), -float(p2), float(F32(3.0) * p2), -float(F32(3.0) * p2)]:
        assert same(ev(["tan", x], protected=True), 1.0)
        assert same(ev(["tan", x], protected=False), F32(_LIBM.tan(x)))
    assert abs(float(ev(["tan", float(p2)], protected=False))) > 1e7  # the pole the protection removes
    assert same(ev(["tan", 0.0]), 0.0)
    q = float(F32(math.pi / 4))
    assert same(ev(["tan", q]), F32(_LIBM.tan(q)))
    # the next float after the protected point is not protected
    nxt = float(np.nextafter(p2, F32(2.0)))
    assert same(ev(["tan", nxt], protected=True), F32(_LIBM.tan(nxt)))


def _gamma_by_hand(x):
    """interpreter_core:127-129 evaluated step by step with the C promotions of a float x:
    1.0f/x is a float; x*sinh(...) is a double product (sinh is the double routine) that SQRT's float
    parameter rounds; x/M_E_F and the product with SQRT are floats; pow is the double routine; its
    double result times SQRT(2.0f*M_PI_F/x) (a float promoted) is rounded once when stored."""
    x = F32(x)
    e_f, pi_f = F32(2.71828174591064), F32(3.14159274101257)
    inv = F32(F32(1.0) / x)
    inner = F32(float(x) * _LIBM.sinh(float(inv)))
    s1 = F32(1.0) if inner < 0 else F32(math.sqrt(float(inner)))
    base = F32(F32(x / e_f) * s1)
    arg2 = F32(F32(F32(2.0) * pi_f) / x)
    s2 = F32(1.0) if arg2 < 0 else F32(math.sqrt(float(arg2)))
    return F32(_LIBM.pow(float(base), float(x)) * float(s2))


@pytest.mark.parametrize("x", [1.0, 2.0, 0.5, 3.5, 7.0])
def test_gamma_promotion_chain(x):
    assert same(ev(["gamma", x]), _gamma_by_hand(x))


@pytest.mark.parametrize("x", [2.0, 3.0, 4.5, 6.0, 10.0])
def test_gamma_approximates_the_gamma_function(x):
    """Windschitl's approximation sqrt(2pi/x) ((x/e) sqrt(x sinh(1/x)))^x of Gamma(x): a dropped square
    root or a wrong exponent misses Gamma by orders of magnitude, the formula itself by < 1e-4."""
    assert float(ev(["gamma", x])) == pytest.approx(math.gamma(x), rel=1e-4)


def test_gamma_signed_zero_and_negative():
    # 1/(+0) = +inf: sinh(inf) = inf, 0*inf = NaN -> pow(NaN, 0) = 1, times SQRT(inf) = inf
    assert same(ev(["gamma", 0.0]), _gamma_by_hand(0.0))
    assert same(ev(["gamma", -0.0]), _gamma_by_hand(-0.0))
    assert same(ev(["gamma", -1.5], protected=True), _gamma_by_hand(-1.5))


# --- IF_THEN_ELSE (interpreter_core:1-4) --------------------------------------------------------

@pytest.mark.parametrize("c,want", [(1.0, 2.0), (0.0, 3.0), (-0.0, 3.0), (NAN, 2.0), (-7.5, 2.0), (1e-40, 2.0)])
def test_if_then_else_selects_by_the_first_child(c, want):
    assert same(ev(["ite", c, 2.0, 3.0]), want)


def test_if_then_else_inside_expression():
    # ite (x0 > x1) (x0 - x1) (x1 - x0) = |x0 - x1|
    prog = ["ite", ">", "x0", "x1", "-", "x0", "x1", "-", "x1", "x0"]
    for a, b in [(3.0, 1.0), (1.0, 3.0), (2.0, 2.0)]:
        assert same(ev(prog, row=(a, b, 0, 0, 0)), abs(a - b))


# --- terminals (interpreter_core:152-793, functions.h:19-57) ------------------------------------

MATH_VALUE = {"pi": math.pi, "pi_2": math.pi / 2, "pi_4": math.pi / 4, "1_pi": 1 / math.pi, "2_pi": 2 / math.pi,
              "2_sqrtpi": 2 / math.sqrt(math.pi), "sqrt2": math.sqrt(2), "sqrt1_2": math.sqrt(0.5), "e": math.e,
              "log2e": 1 / math.log(2), "log10e": 1 / math.log(10), "ln2": math.log(2), "ln10": math.log(10)}


@pytest.mark.parametrize("name", NAMED)
def test_named_constants_are_the_nearest_floats(name):
    """Each functions.h literal (M_PI_F, ...) is the float nearest the named constant."""
    assert same(ev([name]), F32(MATH_VALUE[name]))


@pytest.mark.parametrize("tok,val", [("0", 0.0), ("1", 1.0), ("2", 2.0), ("7", 7.0), ("100", 100.0),
                                     ("n1", -1.0), ("n7", -7.0), ("n100", -100.0)])
def test_integer_terminals(tok, val):
    assert same(ev([tok]), val)


def test_attribute_and_constant_terminals():
    assert same(ev(["x3"], row=(0, 0, 0, 6.25, 0)), 6.25)
    assert same(ev([-4.75]), -4.75)
    assert same(ev(["-", "x2", "x0"], row=(1.0, 0, 10.0, 0, 0)), 9.0)


# --- oracle_exact_sums ---------------------------------------------------------------------------

def test_exact_sums_equal_float_sums_of_integer_errors():
    """With integer-valued errors below 2^24 every float sum is exact, so the exact (double) sum
    equals the sequential float sum and numpy's count."""
    b = Builder(4)
    data = np.zeros((3001, 5), np.float32)
    rng = np.random.default_rng(3)
    data[:, 0] = rng.integers(-50, 50, 3001)
    data[:, 4] = rng.integers(-50, 50, 3001)
    pop = b.population([["x0"], ["+", "x0", 1.0], ["*", 2.0, "x0"]])
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS)
    sums, bad = oracle.exact_sums(cfg, b.table, pop, data)
    x, y = data[:, 0].astype(np.int64), data[:, 4].astype(np.int64)
    want = [np.abs(x - y).sum(), np.abs(x + 1 - y).sum(), np.abs(2 * x - y).sum()]
    assert list(sums) == [float(v) for v in want]
    assert list(bad) == [0, 0, 0]
    vec, _ = oracle.interpret(cfg, b.table, pop, data, alpha=0.0)
    assert np.array_equal(vec, (np.array(want, np.float64) / 3001).astype(np.float32))


def test_exact_sums_are_exact_on_float_errors():
    """The sum of the float per-row errors, exactly (math.fsum), not their sequential float sum."""
    b = Builder(4)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 11, 5000, 4)
    pop = b.population([["*", "x0", "x1"], ["sin", "x2"]])
    for ef in (ppiamd.EF_ABS, ppiamd.EF_SQUARED):
        cfg = oracle.make_cfg(ppiamd.SEQ, ef)
        sums, bad = oracle.exact_sums(cfg, b.table, pop, data)
        preds = [(data[:, 0] * data[:, 1]).astype(np.float32), np.sin(data[:, 2].astype(np.float64)).astype(np.float32)]
        for k, p in enumerate(preds):
            d = (p - data[:, 4]).astype(np.float32)
            e = np.abs(d) if ef == ppiamd.EF_ABS else (d * d).astype(np.float32)
            assert sums[k] == pytest.approx(math.fsum(e.astype(np.float64)), rel=1e-15)
        assert list(bad) == [0, 0]


def test_exact_sums_reducemax_and_bad_rows():
    b = Builder(4)
    data = np.zeros((6, 5), np.float32)
    data[:, 0] = [1, 2, 3, 4, 5, 6]
    pop = b.population([["x0"], ["exp", "exp", "x0"], ["x0"]])
    pop.size[2] = 0
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_ABS, reducemax=True)
    sums, bad = oracle.exact_sums(cfg, b.table, pop, data)
    assert sums[0] == 6.0 * 6  # max(err * nlin)
    assert list(bad) == [0, 1, 1]  # exp(exp(5)) overflows; size 0 is bad


# --- float-computation gradient mode -------------------------------------------------------------

def test_float_computation_gradient_closed_forms():
    """oracle_grad_mag_f64(f32_values=1): operator values are the float results, tangents exact at
    those float operands.  For c*x0 and sin(c*x0) under the squared error the gradient is
    mean(2 (pred - y) x0) and mean(2 (pred - y) cos(c x0) x0) with float pred and float c*x0."""
    b = Builder(4)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_ADJOINT, 5, 700, 5)[:, [0, 1, 2, 3, 5]]
    c = 0.7
    pop = b.population([["*", c, "x0"], ["sin", "*", c, "x0"], ["+", "x1", c]])
    cfg = oracle.make_cfg(ppiamd.SEQ, ppiamd.EF_SQUARED)
    g, m = oracle.grad_mag_f64(cfg, b.table, pop, data, max_consts=2, f32_values=True)
    x0, x1, y = data[:, 0], data[:, 1], data[:, 4].astype(np.float64)
    cx = (F32(c) * x0).astype(np.float32)
    p0 = cx.astype(np.float64)
    p1 = np.sin(p0).astype(np.float32).astype(np.float64)
    p2 = (x1 + F32(c)).astype(np.float32).astype(np.float64)
    want0 = np.mean(2 * (p0 - y) * x0.astype(np.float64))
    want1 = np.mean(2 * (p1 - y) * np.cos(p0) * x0.astype(np.float64))
    want2 = np.mean(2 * (p2 - y))
    assert g[0, 0] == pytest.approx(want0, rel=1e-12)
    assert g[1, 0] == pytest.approx(want1, rel=1e-12)
    assert g[2, 0] == pytest.approx(want2, rel=1e-12)
    assert m[0, 0] == pytest.approx(np.mean(np.abs(2 * (p0 - y) * x0.astype(np.float64))), rel=1e-12)
    # the FP64 mode differs from it (the forward values are doubles there)
    g64, _ = oracle.grad_mag_f64(cfg, b.table, pop, data, max_consts=2, f32_values=False)
    assert g64[1, 0] != g[1, 0]


def test_math_hard_inputs_fixture():
    """The hard inputs of tests/golden/math_hard_inputs.npz are what their generator says: glibc's double
    result (Python's math module calls the C library) lies within 256 ulps of a float rounding boundary,
    and neighbouring inputs that are not listed do not."""
    import math
    import os
    import struct

    data = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "math_hard_inputs.npz"))
    assert sorted(data.files) == sorted(["exp", "log", "sin", "cos", "tan", "exp10", "exp2", "log10", "log2"])

    def hard(y):
        if not (2.0 ** -126 <= abs(y) < 2.0 ** 128):
            return False
        d = (struct.unpack("<Q", struct.pack("<d", y))[0] & 0x1fffffff) - 0x10000000
        return -256 <= d <= 256

    fns = {"exp": math.exp, "log": math.log, "sin": math.sin, "cos": math.cos, "tan": math.tan,
           "log10": math.log10, "log2": math.log2}
    for name, arr in data.items():
        assert arr.dtype == np.uint32 and arr.size > 400 and np.all(np.diff(arr.astype(np.int64)) > 0), name
        x = arr.view(np.float32)
        assert np.all(np.isfinite(x)), name
        if name not in fns:
            continue
        listed = set(arr.tolist())
        for bits in arr[:: max(1, arr.size // 64)].tolist():
            xf = float(np.uint32(bits).view(np.float32))
            assert hard(fns[name](xf)), (name, hex(bits))
            nb = bits + 1
            if nb not in listed:
                xn = float(np.uint32(nb).view(np.float32))
                if math.isfinite(xn) and (name not in ("log", "log10", "log2") or xn > 0):
                    assert not hard(fns[name](xn)), (name, hex(nb))
