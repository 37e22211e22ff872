"""Symbol numbering (reference script/read_grammar.py) against vectors the reference generator wrote."""
import json
import os

import pytest

import ppiamd
from ppiamd import _grammar as grammar

HERE = os.path.dirname(os.path.abspath(__file__))
VECTORS = json.load(open(os.path.join(HERE, "golden", "symbol_vectors.json")))


@pytest.mark.parametrize("key", sorted(VECTORS))
def test_numbering_matches_reference_generator(key):
    v = VECTORS[key]
    g = grammar.parse_bnf(v["bnf"])
    assert g.values() == v["values"]
    assert g.max_rule == v["max_rule"]
    assert (not g.uses_const) == v["not_using_t_const"]


def test_engine_table_maps_every_evaluable_terminal():
    g = grammar.parse_bnf(VECTORS["mixed_order"]["bnf"])
    vals = g.values()
    table = {s: (op, x) for s, op, x in g.engine_table()}
    assert table[vals["T_SUB"]][0] == grammar.OPS["SUB"]
    assert table[vals["T_IF_THEN_ELSE"]][0] == grammar.OPS["IF_THEN_ELSE"]
    assert table[vals["T_PI"]] == (grammar.OPS["VALUE"], pytest.approx(3.14159274101257))
    assert table[vals["T_N4"]] == (grammar.OPS["VALUE"], -4.0)
    assert table[vals["T_100"]] == (grammar.OPS["VALUE"], 100.0)
    assert table[vals["T_CONST"]][0] == grammar.OPS["CONST"]
    assert table[vals["T_ATTRIBUTE"]][0] == grammar.OPS["ATTRIBUTE"]
    # attrK symbols are rewritten to T_ATTRIBUTE by PPI's decoder; they are not evaluable themselves.
    assert vals["T_ATTR7"] not in table


def test_symbol_header_declares_the_same_enum(tmp_path):
    g = grammar.parse_bnf(VECTORS["no_const"]["bnf"])
    text = g.symbol_header()
    assert "NOT_USING_T_CONST" in text and "T_1 = TERMINAL_MIN" in text and "T_ATTR0 = ATTRIBUTE_MIN + 0" in text
    # The CLI writes the header and the engine table include.
    src = tmp_path / "g.bnf"
    src.write_text(VECTORS["config_regression"]["bnf"])
    assert grammar.main(["-g", str(src), "-o", str(tmp_path / "out")]) == 0
    inc = (tmp_path / "out" / "ppi_symbol_table.inc").read_text()
    assert "{10000, 20, 0.000000000e+00f}" in inc  # T_ADD -> ADD


def test_terminal_op_integer_and_named_constants():
    assert grammar.terminal_op("T_0") == (grammar.OPS["VALUE"], 0.0)
    assert grammar.terminal_op("T_N100") == (grammar.OPS["VALUE"], -100.0)
    assert grammar.terminal_op("T_101") is None
    assert grammar.terminal_op("T_N0") is None
    assert grammar.terminal_op("T_LN10")[0] == grammar.OPS["VALUE"]
    assert grammar.terminal_op("T_FOO") is None
    assert ppiamd.OPS["TAN"] == 64


def test_rule_tables_match_the_reference_generator_output():
    """The decoder's tables (gramatica[] / tamanhos[], ppi.cc:106-182 reads them) equal what the
    reference's generator wrote for problem/teste/grammar_bnf1 (its checked-in problem/teste/grammar)."""
    gv = json.load(open(os.path.join(HERE, "golden", "grammar_rules.json")))
    g = grammar.parse_bnf(gv["bnf"])
    nts, alts = g.rule_tables()
    name = {v: k for k, v in g.values().items()}
    assert [name[n] for n in nts] == gv["nonterminals"]
    assert nts == list(range(len(nts)))  # gramatica[] is indexed by the non-terminal's enum value
    assert [[[name[s] for s in alt] for alt in rules] for rules in alts] == gv["rules"]
    assert g.max_rule == gv["max_rule"]
