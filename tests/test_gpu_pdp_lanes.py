"""GPU: the PDP strategy's lane-dependent sentinels.

Under the PDP strategy the reference (accelerator.cl:173-268) gives each of its
pdp_lanes work-items the rows lane, lane + pdp_lanes, ...; a work-item that meets a
non-finite error stores MAXFLOAT and stops, and the work-items' partials meet in a
tree.  Whether a program scores MAXFLOAT or MAXFLOAT/nlin then depends on the lane
of an overflowing row (REDUCEMAX: an infinite error*nlin survives only in a lane
with no bad row) and on the tree's partial sums next to MAXFLOAT.  Each program
below puts an infinite error and huge finite ones in chosen lanes; the engine must
give the oracle's bits in every form, with REDUCEMAX on and off, plain and batched.
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


def lane_data(nlin=NLIN):
    """Columns x0..x9 (ones but for the listed rows) and a zero target."""
    cols = np.ones((max(nlin, NLIN), 11), dtype=np.float32)
    cols[:, 10] = 0.0
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


def oracle_vec(b, pop, data, reducemax):
    dpl, ppl, bpl = oracle.reference_geometry(data.shape[0], pop.n, mls=LANES)
    cfg = oracle.make_cfg(ppiamd.PDP, ppiamd.EF_SQUARED, reducemax, False, None, dpl, ppl, bpl)
    return oracle.interpret(cfg, b.table, pop, data, alpha=ALPHA, best_size=1)[0]


@pytest.mark.parametrize("reducemax", [False, True])
@pytest.mark.parametrize("form", [ppiamd.FORM_FP, ppiamd.FORM_PDP, ppiamd.FORM_PP])
def test_pdp_lane_sentinels(eng, form, reducemax):
    b = Builder(10)
    pop = b.population(PROGRAMS)
    data = lane_data()
    print(f"case pdp-lanes form={form} reducemax={reducemax}", flush=True)
    eng.configure(b.table, strategy=ppiamd.PDP, error_fn=ppiamd.EF_SQUARED, reducemax=reducemax, protected=False,
                  form=form, max_local_size=LANES, compute_units=256)
    eng.set_dataset(data)
    vec, _ = eng.evaluate(pop, alpha=ALPHA, best_size=1)
    ovec = oracle_vec(b, pop, data, reducemax)
    same = agree(vec, ovec, data.shape[0])
    assert same.all(), f"programs {np.nonzero(~same)[0]}: {vec[~same]} vs {ovec[~same]}"


@pytest.mark.parametrize("reducemax", [False, True])
def test_pdp_lane_sentinels_batched(reducemax):
    b = Builder(10)
    pop = b.population(PROGRAMS)
    datas = [lane_data(), lane_data(700)]
    print(f"case pdp-lanes batched reducemax={reducemax}", flush=True)
    bt = ppiamd.Batch(0)
    bt.configure(b.table, strategy=ppiamd.PDP, error_fn=ppiamd.EF_SQUARED, reducemax=reducemax, protected=False,
                 max_local_size=LANES, compute_units=256)
    ids = [bt.add_population(bt.add_dataset(d), pop) for d in datas]
    bt.build()
    for rep in range(2):  # graph replays
        bt.run()
        for k, d in enumerate(datas):
            vec, _ = bt.results(ids[k], alpha=ALPHA)
            ovec = oracle_vec(b, pop, d, reducemax)
            same = agree(vec, ovec, d.shape[0])
            assert same.all(), f"rep {rep} pair {k} programs {np.nonzero(~same)[0]}: {vec[~same]} vs {ovec[~same]}"
    bt.close()
