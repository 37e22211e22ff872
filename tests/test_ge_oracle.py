"""Pins the oracle's restatement of PPI's grammatical-evolution loop (oracle/ppi_ge_oracle.c) on the CPU.

* the random number generator against draws of the reference's own XorShift128Plus, compiled from
  src/util/Random.h (tests/golden/rng_vectors.json, written by tests/golden/make_rng_vectors.py);
* decode (ppi.cc:106-182) against hand-derived genomes of a small grammar: rule choice by the gene's
  value modulo the rule count, least significant bit first, T_CONST values, attribute rewriting,
  running out of bits and the phenotype-size limit;
* breeding (ppi.cc:422-452,680-814) through invariants a slip would break (tournament direction,
  crossover segment structure, clones) and a step traced by hand from the reference generator's
  own draws.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
import ppiamd
from ppiamd import _grammar as grammar

HERE = os.path.dirname(os.path.abspath(__file__))
RNG = json.load(open(os.path.join(HERE, "golden", "rng_vectors.json")))


@pytest.mark.parametrize("seed", sorted(RNG["streams"], key=int))
def test_rng_matches_the_reference_generator(seed):
    v = RNG["streams"][seed]
    ints, reals = oracle.rng_draws(int(seed), RNG["n"])
    assert [str(int(x)) for x in ints] == v["int"]
    assert list(reals) == [float(x) for x in v["real"]]
    # Probability(0.66) is Real() < 0.66 on its own stream
    assert [int(r < 0.66) for r in reals] == v["p066"]


def test_rng_real_is_the_int_over_two_to_the_64():
    ints, reals = oracle.rng_draws(99, 50)
    for i, r in zip(ints, reals):
        assert r == float(int(i)) / 2.0 ** 64  # double(Int()) / 2^64, one rounding
        assert 0.0 <= r < 1.0


# --- decode ---------------------------------------------------------------------------------

BNF = ("S = <e>\n"
       "P = <e> ::= <b> <e> <e> | <u> <e> | <t>\n"
       "    <b> ::= + | - | *\n"
       "    <u> ::= sin | neg\n"
       "    <t> ::= attr0 | attr3 | const | 2\n")


def setup(**kw):
    g = grammar.parse_bnf(BNF)
    return g, ppiamd.GeSetup(g, **kw)


def bits(values, width):
    """Genes least significant bit first (decode_rule, ppi.cc:118-121)."""
    out = []
    for v in values:
        out += [(v >> i) & 1 for i in range(width)]
    return out


def test_effective_bits_and_size():
    g, ge = setup(number_of_bits=400, bits_per_gene=2)  # <e> and <b> have 3 rules, <t> 4: 2 bits suffice
    assert oracle.ge_effective(ge) == (2, g.max_rule * 400 // 2)
    big = grammar.parse_bnf("S = <e>\nP = <e> ::= " + " | ".join(f"attr{k}" for k in range(300)) + "\n")
    gb = ppiamd.GeSetup(big, number_of_bits=100, bits_per_gene=8)
    assert oracle.ge_effective(gb)[0] == 9  # ceil(log2(300)) = 9 > 8 (ppi.cc:258-264)
    _, capped = setup(number_of_bits=400, bits_per_gene=8, max_phenotype=7)
    assert oracle.ge_effective(capped)[1] == 7


def test_decode_hand_derivation():
    g, ge = setup(number_of_bits=8 * 12, bits_per_gene=8, bits_per_constant=8, min_constant=-1.0, max_constant=3.0)
    v = g.values()
    # S -> e; e -> rule 0 (b e e); b -> '-' (1); e -> rule 2 (t); t -> attr3 (1);
    # e -> rule 1 + 3k (u e): value 4 = 1 mod 3; u -> neg (1 + 2k = 3); e -> t (rule 2: 5); t -> const (2 + 4k = 6),
    # then 8 bits of constant = 51 -> -1 + 51 * 4 / 255 = -0.2
    genes = [0, 0, 1, 2, 1, 4, 3, 5, 6, 51, 0, 0]
    genome = np.array([bits(genes, 8)], dtype=np.uint8)
    ph, ep, sz, mx = oracle.ge_decode(ge, genome)
    want = [v["T_SUB"], v["T_ATTRIBUTE"], v["T_NEG"], v["T_CONST"]]
    assert sz[0] == 4 and list(ph[0, :4]) == want
    assert ep[0, 1] == 3.0  # T_ATTR3 -> T_ATTRIBUTE, ephemeral 3 (ppi.cc:167-171)
    lo, hi = np.float32(-1.0), np.float32(3.0)
    expect = np.float32(float(lo) + float(np.float32(np.float32(51) * (hi - lo))) / 255.0)
    assert ep[0, 3] == expect and abs(float(expect) + 0.2) < 1e-6


def test_decode_runs_out_of_bits_and_size_limit():
    g, ge = setup(number_of_bits=8 * 3, bits_per_gene=8)
    # S -> e (0), e -> b e e (0), b -> '+' (0), then no bits left for the first operand: killed
    ph, ep, sz, _ = oracle.ge_decode(ge, np.array([bits([0, 0, 0], 8)], np.uint8))
    assert sz[0] == 0
    # a constant without enough bits for its value decodes to 0 and consumes nothing (ppi.cc:124-127)
    g, ge = setup(number_of_bits=8 * 3 + 4, bits_per_gene=8, bits_per_constant=8)
    ph, ep, sz, _ = oracle.ge_decode(ge, np.array([bits([0, 2, 2], 8) + [1, 1, 1, 1]], np.uint8))
    assert sz[0] == 1 and ph[0, 0] == g.values()["T_CONST"] and ep[0, 0] == 0.0
    # -mps: a rule decoded at position >= the size limit kills the individual (ppi.cc:143)
    g, ge = setup(number_of_bits=8 * 12, bits_per_gene=8, max_phenotype=3)
    ph, ep, sz, _ = oracle.ge_decode(ge, np.array([bits([0, 0, 1, 2, 1, 4, 3, 5, 6, 51, 0, 0], 8)], np.uint8))
    assert sz[0] == 0


def test_decode_round_trip_of_random_derivations():
    """Genes chosen to realise a random derivation (value = rule index + k * rule count for random k)
    decode back to that derivation's prefix symbols."""
    g, ge = setup(number_of_bits=8 * 200, bits_per_gene=8, bits_per_constant=8)
    v = g.values()
    nts, alts = g.rule_tables()
    rng = np.random.default_rng(5)
    for trial in range(40):
        genes, want = [], []

        def expand(nt, depth):
            rules = alts[nt]
            r = int(rng.integers(len(rules))) if depth < 5 else len(rules) - 1
            if nt == v["NT_E"] and depth >= 5:
                r = 2
            genes.append(r + len(rules) * int(rng.integers(0, 256 // len(rules))))
            for s in rules[r]:
                if s >= 10000:
                    want.append(v["T_ATTRIBUTE"] if s >= 20000 else s)
                    if s == v["T_CONST"]:
                        genes.append(int(rng.integers(256)))
                else:
                    expand(s, depth + 1)
        expand(0, 0)
        genes += [0] * (200 - len(genes))
        ph, ep, sz, _ = oracle.ge_decode(ge, np.array([bits(genes, 8)], np.uint8))
        assert sz[0] == len(want) and list(ph[0, :sz[0]]) == want, trial


# --- breeding -------------------------------------------------------------------------------

def test_tournament_picks_the_lowest_fitness():
    """With tournaments far larger than the population, no crossover and no mutation, every child is
    a clone of the best parent (ppi.cc:428-441: a competitor replaces the winner only when lower)."""
    g, ge = setup(population_size=8, number_of_bits=64, tournament_size=300, crossover_rate=0.0,
                  mutation_rate=0.0, seed=11)
    rng = np.random.default_rng(1)
    parents = rng.integers(0, 2, (8, 64)).astype(np.uint8)
    fit = np.array([5, 3, 9, 1.5, 7, 2, 8, 4], np.float32)
    kids = oracle.ge_breed(ge, parents, fit)
    assert (kids == parents[3]).all()
    kids = oracle.ge_breed(ge, parents, fit, n_immigrants=1)  # slot 0 is the elite's / an immigrant's
    assert (kids[0] == 0).all() and (kids[1:] == parents[3]).all()


def test_crossover_structure():
    """Crossover only (cp = 1, mr = 0) of an all-zeros and an all-ones parent: the two children are
    complements; a two-point cut lies on gene boundaries, a one-point cut anywhere (ppi.cc:686-724)."""
    g, ge = setup(population_size=64, number_of_bits=96, bits_per_gene=8, tournament_size=1,
                  crossover_rate=1.0, mutation_rate=0.0, seed=5)
    parents = np.zeros((64, 96), np.uint8)
    parents[1::2] = 1
    kids = oracle.ge_breed(ge, parents, np.zeros(64, np.float32))
    one = two = 0
    for i in range(0, 64, 2):
        a, b = kids[i].astype(int), kids[i + 1].astype(int)
        cuts = np.nonzero(np.diff(a))[0] + 1
        if (a == b).all():
            continue  # father and mother were the same genome
        assert ((a + b) == 1).all()
        assert len(cuts) <= 2
        if len(cuts) == 2:
            assert all(c % 8 == 0 for c in cuts)  # two-point: gene boundaries
            two += 1
        elif len(cuts) == 1 and cuts[0] % 8:
            one += 1  # a cut inside a gene: one-point
    assert two > 5 and one > 2


def test_breeding_step_traced_from_the_reference_draws():
    """P = 2, one stream, tournaments of one, crossover always, no mutation: the step consumes, in
    order, father = int(r0 * 2), mother = int(r1 * 2), the crossover test r2 < 1, Probability(0.66)
    on r3, the cut draws, and one draw per child for the number of mutated bits (0 of 0 + 1).
    Expected children follow from the reference generator's own draws (rng_vectors.json)."""
    seed = 12345
    reals = [float(x) for x in RNG["streams"][str(seed)]["real"]]
    nb, bg = 32, 8
    g, ge = setup(population_size=2, number_of_bits=nb, bits_per_gene=bg, tournament_size=1,
                  crossover_rate=1.0, mutation_rate=0.0, seed=seed)
    rng = np.random.default_rng(3)
    parents = rng.integers(0, 2, (2, nb)).astype(np.uint8)
    kids = oracle.ge_breed(ge, parents, np.zeros(2, np.float32))
    f, m = parents[int(reals[0] * 2)], parents[int(reals[1] * 2)]
    assert reals[2] < 1.0
    if reals[3] < 0.66:
        a, b = sorted([int(reals[4] * nb) // bg * bg, int(reals[5] * nb) // bg * bg])
        c1 = np.concatenate([f[:a], m[a:b], f[b:]])
        c2 = np.concatenate([m[:a], f[a:b], m[b:]])
    else:
        pt = int(reals[4] * nb)
        c1 = np.concatenate([f[:pt], m[pt:]])
        c2 = np.concatenate([m[:pt], f[pt:]])
    assert (kids[0] == c1).all() and (kids[1] == c2).all()


def test_shrink_and_bitflip_keep_the_genome_length_and_parents_bits():
    """Mutation only (cp = 0): a child differs from its clone either in at most max_bits positions
    (bit flips) or by a left shift of a gene-aligned tail (shrink: memmove, the last bits unchanged)."""
    g, ge = setup(population_size=200, number_of_bits=128, bits_per_gene=8, tournament_size=1,
                  crossover_rate=0.0, mutation_rate=0.05, seed=9)
    rng = np.random.default_rng(4)
    parents = rng.integers(0, 2, (200, 128)).astype(np.uint8)
    kids = oracle.ge_breed(ge, parents, np.zeros(200, np.float32))
    max_bits = math.ceil(np.float32(0.05) * np.float32(128))
    flips = shrinks = 0
    for k in kids:
        d = [int((k != p).sum()) for p in parents]
        src = parents[int(np.argmin(d))]
        if min(d) <= max_bits:
            flips += 1
            continue
        # shrink: k = src[:s] + src[e:] + src[len - (e - s):]
        found = False
        for s in range(0, 128, 8):
            for e in range(s + 8, 129, 8):
                for p in parents:
                    cand = np.concatenate([p[:s], p[e:], p[128 - (e - s):]])
                    if (cand == k).all():
                        found = True
                        break
                if found:
                    break
            if found:
                break
        assert found
        shrinks += 1
    assert flips > 50 and shrinks > 5
