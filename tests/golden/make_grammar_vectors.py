#!/usr/bin/env python3
"""Writes grammar_rules.json: the decoder's rule tables of one reference problem, taken from the
reference's checked-in output of its grammar generator (problem/teste/grammar, the t_rule /
gramatica[] tables that script/read_grammar.py wrote for problem/teste/grammar_bnf1).

Nothing is executed: the generated header is parsed as data (rule names, quantities and symbol
names), and the BNF it was generated from is kept beside it as the input.  Run in a container that
has the reference checked out at /root/reference.
"""
import json
import os
import re

REF = "/root/reference/problem/teste"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    bnf = open(os.path.join(REF, "grammar_bnf1")).read()
    text = open(os.path.join(REF, "grammar")).read()
    max_rule = int(re.search(r"#define MAX_QUANT_SIMBOLOS_POR_REGRA (\d+)", text).group(1))
    rules = {}
    for name, qty, syms in re.findall(r"t_rule (\w+) = \{ (\d+), \{([^}]*)\} \};", text):
        rules[name] = [s.strip() for s in syms.split(",")][: int(qty)]
    heads = re.search(r"t_rule\*\* gramatica\[\] = \{([^}]*)\};", text).group(1)
    order = [h.strip() for h in heads.split(",") if h.strip()]
    tables = []
    for arr in order:
        members = re.search(r"t_rule\* " + arr + r"\[\] = \{([^}]*)\};", text).group(1)
        tables.append([rules[m.strip().lstrip("&")] for m in members.split(",") if m.strip()])
    out = {"source": "problem/teste/grammar (generated from problem/teste/grammar_bnf1)", "bnf": bnf,
           "max_rule": max_rule, "nonterminals": ["NT_" + a[len("regras_"):] for a in order], "rules": tables}
    with open(os.path.join(HERE, "grammar_rules.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
