"""GPU: grammatical evolution on the device (ppi_ge_*) and the device epilogue, against the oracle.

* ppi_engine_evaluate_ranked: alpha*size and the strategy's ranking by kernels give the host
  epilogue's vector and indices bit for bit (every strategy, FP and PDP forms);
* the generated population (generation 0) equals the oracle's ppi_generate_population genome for
  genome, and its decode (ppi.cc:106-182) the oracle's phenotypes;
* the device compiler's tape equals the host compiler's for the same phenotypes, instruction for
  instruction (folded constants included);
* whole runs of ppi_evolve's loop (elitism, tournaments, crossover, mutation, odd populations, one or
  several random streams) follow the oracle's run: every genome, every fitness and the best fitness of
  every generation, bit for bit, with the evaluation in the sequential path's (SEQ, ordered form) and
  evaluate_pp's (PP, ordered form) semantics;
* immigrants start the next population, ahead of the elite.
"""
import numpy as np
import pytest

import oracle
import ppiamd
from ppiamd import _grammar as grammar

pytestmark = pytest.mark.gpu

ALPHA = np.float32(1e-6)

BNF = ("S = <e>\n"
       "P = <e> ::= <b> <e> <e> | <u> <e> | <t> | if_then_else <e> <e> <e>\n"
       "    <b> ::= + | - | * | / | max | <\n"
       "    <u> ::= sin | exp | log | sqrt | neg\n"
       "    <t> ::= attr0 | attr1 | attr2 | attr3 | const | PI | 2\n")


@pytest.fixture(scope="module")
def eng():
    if ppiamd.device_count() == 0:
        pytest.fail("no HIP device visible")
    e = ppiamd.Engine(0)
    yield e
    e.close()


def same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


@pytest.mark.parametrize("strategy", [ppiamd.SEQ, ppiamd.PP, ppiamd.FP, ppiamd.PDP])
@pytest.mark.parametrize("form", [ppiamd.FORM_FP, ppiamd.FORM_PDP])
@pytest.mark.parametrize("best_size", [1, 5])
def test_ranked_epilogue_matches_the_host_epilogue(eng, strategy, form, best_size):
    w = ppiamd.workload("small", n_programs=2500, nlin=1500)
    print(f"case ranked strategy={strategy} form={form} best_size={best_size}", flush=True)
    eng.configure(w.grammar.table, strategy=strategy, error_fn=w.error_fn, form=form, compute_units=256)
    eng.set_dataset(w.data)
    v1, i1 = eng.evaluate(w.pop, alpha=ALPHA, best_size=best_size)
    v2, i2 = eng.evaluate_ranked(w.pop, alpha=ALPHA, best_size=best_size)
    assert same_bits(v1, v2)
    assert list(i1) == list(i2)
    # ties: many equal errors (constant programs), the reference's tie rules decide
    b = ppiamd.Population(w.pop.phenotype.copy(), w.pop.ephemeral.copy(), w.pop.size.copy(), w.pop.stride)
    b.size[::3] = 0
    v1, i1 = eng.evaluate(b, alpha=0.0, best_size=best_size)
    v2, i2 = eng.evaluate_ranked(b, alpha=0.0, best_size=best_size)
    assert same_bits(v1, v2) and list(i1) == list(i2)


def make_ge(eng, strategy=ppiamd.SEQ, form=ppiamd.FORM_ORDERED, nlin=700, **kw):
    g = grammar.parse_bnf(BNF)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 31, nlin, 4)
    eng.configure(g.engine_table(), strategy=strategy, error_fn=ppiamd.EF_ABS, form=form, compute_units=256)
    eng.set_dataset(data)
    setup = ppiamd.GeSetup(g, **kw)
    return g, data, setup, ppiamd.GE(eng, setup)


def oracle_cfg(strategy, nlin, P):
    dpl, ppl, bpl = oracle.reference_geometry(nlin, P)
    return oracle.make_cfg(strategy, ppiamd.EF_ABS, False, True, None, dpl, ppl, bpl)


def test_generation_zero_and_decode(eng):
    g, data, setup, ge = make_ge(eng, population_size=300, number_of_bits=600, seed=77)
    print("case ge generation 0: 300 x 600 alleles", flush=True)
    ge.initialize()
    genomes, fit = ge.population()
    og, ofit, hist, best = oracle.ge_run(oracle_cfg(ppiamd.SEQ, data.shape[0], 300), g.engine_table(), setup, data, 0)
    assert np.array_equal(genomes, og)
    ph, ep, sz = ge.decoded()
    oph, oep, osz, mx = oracle.ge_decode(setup, genomes)
    assert np.array_equal(sz, osz)
    assert (sz > 0).sum() > 30
    for i in range(300):
        assert np.array_equal(ph[i, :sz[i]], oph[i, :sz[i]]), i
        assert same_bits(ep[i, :sz[i]], oep[i, :sz[i]]), i
    assert same_bits(fit, ofit)
    bf, bg = ge.best()
    assert np.float32(bf) == hist[0] and np.array_equal(bg, best)


@pytest.mark.parametrize("fold", [True, False])
def test_device_compiler_matches_the_host_compiler(eng, fold):
    g = grammar.parse_bnf(BNF)
    data = ppiamd.dataset(ppiamd.DS_FORMULA_SMALL, 32, 500, 4)
    eng.configure(g.engine_table(), strategy=ppiamd.SEQ, error_fn=ppiamd.EF_ABS, form=ppiamd.FORM_PDP,
                  fold_constants=fold)
    eng.set_dataset(data)
    setup = ppiamd.GeSetup(g, population_size=2000, number_of_bits=1200, bits_per_constant=12, seed=5)
    ge = ppiamd.GE(eng, setup)
    print(f"case ge compile vs host fold={fold}", flush=True)
    ge.initialize()
    meta_d, code_d = ge.tape()
    ph, ep, sz = ge.decoded()
    pop = ppiamd.Population(ph.ravel().copy(), ep.ravel().copy(), sz.copy(), ph.shape[1])
    eng.upload(pop)
    meta_h, code_h = eng.resident_tape()
    assert np.array_equal(meta_d, meta_h)
    assert code_d.shape == code_h.shape
    diff = np.nonzero((code_d != code_h).any(axis=1))[0]
    assert diff.size == 0, f"{diff.size} instructions differ, first at {diff[:5]}: {code_d[diff[:3]]} vs {code_h[diff[:3]]}"
    assert (meta_d[:, 1] > 3).sum() > 50  # programs of some length took part
    ge.close()


CASES = [
    dict(strategy=ppiamd.SEQ, population_size=200, number_of_bits=400, elitism=True, threads=1, seed=3),
    dict(strategy=ppiamd.SEQ, population_size=101, number_of_bits=320, elitism=False, threads=3, seed=11,
         mutation_rate=0.05, tournament_size=4),
    dict(strategy=ppiamd.SEQ, population_size=64, number_of_bits=256, elitism=True, threads=5, seed=1234,
         crossover_rate=0.5, mutation_rate=0.2, bits_per_constant=10, max_phenotype=40),
    dict(strategy=ppiamd.PP, population_size=150, number_of_bits=500, elitism=True, threads=2, seed=9),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_evolution_follows_the_oracle(eng, case):
    kw = dict(CASES[case])
    strategy = kw.pop("strategy")
    gens = 6
    g, data, setup, ge = make_ge(eng, strategy=strategy, **kw)
    print(f"case ge run {case}: {CASES[case]}", flush=True)
    ge.initialize()
    hist_d = [ge.best()[0]]
    for _ in range(gens):
        ge.evolve(1)
        hist_d.append(ge.best()[0])
    genomes, fit = ge.population()
    P = kw["population_size"]
    og, ofit, hist, best = oracle.ge_run(oracle_cfg(strategy, data.shape[0], P), g.engine_table(), setup, data, gens)
    assert np.array_equal(genomes, og)
    assert same_bits(fit, ofit)
    assert same_bits(np.array(hist_d, np.float32), hist)
    assert np.array_equal(ge.best()[1], best)
    assert ge.stats().generation == gens


def test_immigrants_start_the_next_population(eng):
    g, data, setup, ge = make_ge(eng, population_size=50, number_of_bits=256, elitism=True, seed=21)
    ge.initialize()
    rng = np.random.default_rng(2)
    imm = rng.integers(0, 2, (3, 256)).astype(np.uint8)
    ge.set_immigrants(imm)
    _, best = ge.best()
    ge.evolve(1)
    genomes, _ = ge.population()
    assert np.array_equal(genomes[:3], imm)
    assert np.array_equal(genomes[3], best)  # the elite follows the immigrants (ppi.cc:882-887)
