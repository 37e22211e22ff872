#!/usr/bin/env python3
"""Writes symbol_vectors.json: Symbol enum values that the reference's own build-time
generator (script/read_grammar.py) assigns for a few BNF grammars.

Run in a container that has the reference checked out at /root/reference.  The
reference script is executed as a child process; only its output (the enum
text of the generated ``symbol`` header, parsed to name -> value) is kept.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

GRAMMARS = {
    "config_regression": (
        "S = <expr>\n"
        "P = <expr> ::= <bin> <expr> <expr> | <un> <expr> | <term>\n"
        "    <bin> ::= + | - | * | /\n"
        "    <un> ::= sin | cos | exp | log | sqrt\n"
        "    <term> ::= attr0 | attr1 | attr2 | attr3 | const\n"),
    "mixed_order": (
        "S = <e> <e>\n"
        "P = <e> ::= <op> <e> <e> | <pre> <e> | <v> | if_then_else <c> <e> <e>\n"
        "    <op> ::= - | + | pow | mean\n"
        "    <pre> ::= sqrt | neg | sin\n"
        "    <c> ::= < <v> <v> | >= <v> 3 | not <c>\n"
        "    <v> ::= attr2 | attr7 | PI | N4 | 100 | const\n"),
    "no_const": (
        "S = <pol>\n"
        "P = <pol> ::= attr0 | 1 | + <pol> <pol> | - <pol> <pol> | * <pol> <pol>\n"),
}


def enum_values(header: str) -> dict:
    body = re.search(r"typedef enum \{(.*)\} Symbol;", header, re.S).group(1)
    consts = {"TERMINAL_MIN": 10000, "ATTRIBUTE_MIN": 20000}
    out, nxt = {}, 0
    for item in [x.strip() for x in body.split(",") if x.strip()]:
        if "=" in item:
            name, expr = [s.strip() for s in item.split("=", 1)]
            val = eval(expr, {}, dict(consts))
        else:
            name, val = item, nxt
        out[name] = val
        nxt = val + 1
    return out


def main():
    vectors = {}
    for key, bnf in GRAMMARS.items():
        with tempfile.TemporaryDirectory() as d:
            g = os.path.join(d, "grammar.bnf")
            open(g, "w").write(bnf)
            subprocess.run([sys.executable, os.path.join(REF, "script", "read_grammar.py"), "-o", d, "-g", g,
                            "-i", os.path.join(REF, "src", "interpreter", "interpreter_core"),
                            "-p", os.path.join(REF, "src", "interpreter", "ppi_individual_print_function")],
                           check=True, capture_output=True)
            header = open(os.path.join(d, "symbol")).read()
            grammar_h = open(os.path.join(d, "grammar")).read()
        maxrule = int(re.search(r"MAX_QUANT_SIMBOLOS_POR_REGRA (\d+)", grammar_h).group(1))
        vectors[key] = {"bnf": bnf, "values": enum_values(header), "not_using_t_const": "NOT_USING_T_CONST" in header,
                        "max_rule": maxrule}
    with open(os.path.join(HERE, "symbol_vectors.json"), "w") as f:
        json.dump(vectors, f, indent=1, sort_keys=True)
    print("wrote", len(vectors), "grammars")


if __name__ == "__main__":
    main()
