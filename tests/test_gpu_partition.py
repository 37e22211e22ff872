"""GPU: results that depend on the reference's partition of the rows.

PDP strategy (accelerator.cl:173-268): each of pdp_lanes work-items takes the rows
lane, lane + pdp_lanes, ...; a work-item that meets a non-finite error stores
MAXFLOAT and stops, and the work-items' partials meet in a tree.  Whether a program
scores MAXFLOAT or MAXFLOAT/nlin then depends on the lane of an overflowing row
(REDUCEMAX: an infinite error*nlin survives only in a lane with no bad row) and on
the tree's partial sums next to MAXFLOAT.

DP strategy (accelerator.cl:88-171, accelerator.cc:773-805): work-groups of dp_local
rows reduce by a tree and the host stops at the first non-finite group result.  A
group sum that overflows gives MAXFLOAT/nlin where an overflow of the host's sum
gives MAXFLOAT, and the REDUCEMAX tree keeps a NaN met as its lower operand (hiding
the upper one, an inf included) and drops a NaN met as its upper operand.

Each program below puts NaN, infinite and huge finite errors in chosen rows; the
engine must give the oracle's bits for every sentinel (and the finite sums within
1e-6) under every strategy and form, with REDUCEMAX on and off, plain and batched.
"""
import numpy as np
import pytest

import oracle
import ppiamd
from helpers import Builder, rel_close

pytestmark = pytest.mark.gpu

ALPHA = np.float32(1e-6)
LANES = 256
NLIN = 2048
NATTR = 12  # x0 .. x11, then the target


def lane_data(nlin=NLIN):
    """Columns x0..x11 (ones but for the listed rows) and a zero target."""
    cols = np.ones((max(nlin, NLIN), NATTR + 1), dtype=np.float32)
    cols[:, NATTR] = 0.0
    big, mid = np.float32(1e18), np.float32(3e15)  # squared: 1e36 (x nlin overflows), 9e30 (< 2^103)
    cols[5, 1] = 0.0                      # x0/x1: inf in lane 5
    cols[5 + 3 * LANES, 2] = big          # lane 5, after its bad row
    cols[9, 3] = big                      # lane 9
    cols[9, 4] = cols[10, 4] = mid        # lanes 9 and 10
    cols[9, 5] = cols[13, 5] = mid        # lanes 9 and 13
    cols[6, 6] = 0.0                      # a second bad lane
    cols[133, 7] = mid                    # lane 133, and lane 5 after its bad row:
    cols[5 + LANES, 7] = big
    cols[5, 8] = big                      # lane 5 overflows, then turns bad at row 5 + LANES:
    cols[5 + LANES, 9] = 0.0
    return np.ascontiguousarray(cols[:nlin])


PROGRAMS = [
    ["/", "x2", "x1"],
    ["/", "x3", "x1"],
    ["/", "x4", "x1"],
    ["/", "x5", "x1"],
    ["*", "/", "x4", "x1", "/", "x0", "x6"],
    ["/", "x3", "x0"],
    ["/", "x0", "x0"],
    ["/", "x7", "x1"],
    ["/", "x8", "x9"],
    ["+", "x4", "x5"],
]

DP_LOCAL = NLIN // 256  # ceil(nlin / compute units) rows per DP work-group


def dp_data(nlin=NLIN):
    """x0 ones; (x[2k+1], x[2k+2]) pairs whose quotient is NaN (0/0), inf (1/0) or huge in chosen rows."""
    L = DP_LOCAL
    cols = np.ones((max(nlin, NLIN), NATTR + 1), dtype=np.float32)
    cols[:, NATTR] = 0.0
    big, near = np.float32(1e18), np.float32(1.6e19)  # squared: 1e36 (x nlin overflows), 2.56e38 (two overflow)

    def nan_at(c, r):
        cols[r, c] = cols[r, c + 1] = 0.0

    def inf_at(c, r):
        cols[r, c + 1] = 0.0

    nan_at(1, 3)                      # x1/x2: NaN in lane 3 of group 0 (the max-tree drops it)
    nan_at(3, L)                      # x3/x4: NaN in lane 0 of group 1 (kept)
    nan_at(5, 2)                      # x5/x6: NaN in lane 2 hides the inf in lane 2 + L/2
    inf_at(5, 2 + L // 2)
    cols[0, 7] = cols[1, 7] = near    # x7/x0: two huge errors in group 0 (the group sum overflows)
    cols[0, 8] = cols[L, 8] = near    # x8/x0: in groups 0 and 1 (the host's sum overflows)
    cols[5, 9] = big                  # x9/x0: error*nlin overflows, no bad row
    # Sums within a rounding of float overflow (the reference's own order decides):
    # x10: four errors of 2^126 (with the ones, just above 2^128: inf in every order);
    # x11: three of 2^126 and one of 2^126 - 2^104 (FLT_MAX exactly, finite in every order).
    for r in (0, 500, 1000, 1500):
        cols[r, 10] = cols[r, 11] = np.float32(2.0 ** 63)
    cols[1500, 11] = np.float32(2.0 ** 63 - 2.0 ** 40)
    return np.ascontiguousarray(cols[:nlin])


DP_PROGRAMS = [
    ["/", "x1", "x2"],
    ["/", "x3", "x4"],
    ["/", "x5", "x6"],
    ["/", "x7", "x0"],
    ["/", "x8", "x0"],
    ["/", "x9", "x0"],
    ["+", "/", "x1", "x2", "/", "x9", "x0"],
    ["/", "x0", "x0"],
    ["/", "x10", "x0"],
    ["/", "x11", "x0"],
]

CASES = {"lanes": (lane_data, PROGRAMS), "groups": (dp_data, DP_PROGRAMS)}
STRATEGIES = [ppiamd.SEQ, ppiamd.PP, ppiamd.FP, ppiamd.PDP]


@pytest.fixture(scope="module")
def eng():
    if ppiamd.device_count() == 0:
        pytest.fail("no HIP device visible")
    e = ppiamd.Engine(0)
    yield e
    e.close()


def agree(vec, ovec, nlin):
    """Sentinels (MAXFLOAT, MAXFLOAT/nlin) bit for bit; the finite sums within 1e-6 (their order differs)."""
    sentinel = ovec >= np.float32(np.finfo(np.float32).max) / np.float32(nlin)
    return (vec.view(np.uint32) == ovec.view(np.uint32)) | (~sentinel & rel_close(vec, ovec, 1e-6))


def oracle_vec(b, pop, data, strategy, reducemax):
    dpl, ppl, bpl = oracle.reference_geometry(data.shape[0], pop.n, mls=LANES)
    cfg = oracle.make_cfg(strategy, ppiamd.EF_SQUARED, reducemax, False, None, dpl, ppl, bpl)
    return oracle.interpret(cfg, b.table, pop, data, alpha=ALPHA, best_size=1)[0]


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("reducemax", [False, True])
@pytest.mark.parametrize("form", [ppiamd.FORM_FP, ppiamd.FORM_PDP, ppiamd.FORM_PP])
@pytest.mark.parametrize("strategy", STRATEGIES)
def test_partition_sentinels(eng, strategy, form, reducemax, case):
    make, programs = CASES[case]
    b = Builder(NATTR)
    pop = b.population(programs)
    data = make()
    print(f"case partition {case} strategy={strategy} form={form} reducemax={reducemax}", flush=True)
    eng.configure(b.table, strategy=strategy, error_fn=ppiamd.EF_SQUARED, reducemax=reducemax, protected=False,
                  form=form, max_local_size=LANES, compute_units=256)
    eng.set_dataset(data)
    vec, _ = eng.evaluate(pop, alpha=ALPHA, best_size=1)
    ovec = oracle_vec(b, pop, data, strategy, reducemax)
    same = agree(vec, ovec, data.shape[0])
    assert same.all(), f"programs {np.nonzero(~same)[0]}: {vec[~same]} vs {ovec[~same]}"


@pytest.mark.parametrize("reducemax", [False, True])
@pytest.mark.parametrize("strategy", STRATEGIES)
def test_partition_sentinels_batched(strategy, reducemax):
    print(f"case partition batched strategy={strategy} reducemax={reducemax}", flush=True)
    b = Builder(NATTR)
    for case in sorted(CASES):
        make, programs = CASES[case]
        pop = b.population(programs)
        datas = [make(), make(700)]
        bt = ppiamd.Batch(0)
        bt.configure(b.table, strategy=strategy, error_fn=ppiamd.EF_SQUARED, reducemax=reducemax, protected=False,
                     max_local_size=LANES, compute_units=256)
        ids = [bt.add_population(bt.add_dataset(d), pop) for d in datas]
        bt.build()
        for rep in range(2):  # graph replays
            bt.run()
            for k, d in enumerate(datas):
                vec, _ = bt.results(ids[k], alpha=ALPHA)
                ovec = oracle_vec(b, pop, d, strategy, reducemax)
                same = agree(vec, ovec, d.shape[0])
                assert same.all(), f"{case} rep {rep} pair {k} programs {np.nonzero(~same)[0]}: {vec[~same]} vs {ovec[~same]}"
        bt.close()


# From 8,192 rows on, a program whose result needs all its rows in the reference's partition is
# queued on the recheck's deep list (row errors computed across the device, then one workgroup
# per program); the list holds 16 programs and the rest are recomputed in the recheck kernel.
# Every program below appears three times, so both paths run.
LONG = 3 * 8192 + 5


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("reducemax", [False, True])
@pytest.mark.parametrize("form", [ppiamd.FORM_FP, ppiamd.FORM_PDP, ppiamd.FORM_PP])
@pytest.mark.parametrize("strategy", STRATEGIES)
def test_partition_sentinels_long(eng, strategy, form, reducemax, case):
    make, programs = CASES[case]
    b = Builder(NATTR)
    pop = b.population(programs * 3)
    data = make(LONG)
    print(f"case partition long {case} strategy={strategy} form={form} reducemax={reducemax}", flush=True)
    eng.configure(b.table, strategy=strategy, error_fn=ppiamd.EF_SQUARED, reducemax=reducemax, protected=False,
                  form=form, max_local_size=LANES, compute_units=256)
    eng.set_dataset(data)
    vec, _ = eng.evaluate(pop, alpha=ALPHA, best_size=1)
    ovec = oracle_vec(b, pop, data, strategy, reducemax)
    same = agree(vec, ovec, data.shape[0])
    assert same.all(), f"programs {np.nonzero(~same)[0]}: {vec[~same]} vs {ovec[~same]}"


@pytest.mark.parametrize("reducemax", [False, True])
@pytest.mark.parametrize("strategy", STRATEGIES)
def test_partition_sentinels_long_batched(strategy, reducemax):
    print(f"case partition long batched strategy={strategy} reducemax={reducemax}", flush=True)
    b = Builder(NATTR)
    for case in sorted(CASES):
        make, programs = CASES[case]
        pop = b.population(programs * 3)
        datas = [make(LONG), make(700), make(LONG + 2048)]
        bt = ppiamd.Batch(0)
        bt.configure(b.table, strategy=strategy, error_fn=ppiamd.EF_SQUARED, reducemax=reducemax, protected=False,
                     max_local_size=LANES, compute_units=256)
        ids = [bt.add_population(bt.add_dataset(d), pop) for d in datas]
        bt.build()
        for rep in range(2):  # graph replays
            bt.run()
            for k, d in enumerate(datas):
                vec, _ = bt.results(ids[k], alpha=ALPHA)
                ovec = oracle_vec(b, pop, d, strategy, reducemax)
                same = agree(vec, ovec, d.shape[0])
                assert same.all(), f"{case} rep {rep} pair {k} programs {np.nonzero(~same)[0]}: {vec[~same]} vs {ovec[~same]}"
        bt.close()
