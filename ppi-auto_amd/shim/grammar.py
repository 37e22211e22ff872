"""BNF grammar -> Symbol numbering, symbol header and engine symbol table.

PPI compiles each problem against a generated ``symbol`` header whose enum
values are the phenotype's opcodes (reference: script/read_grammar.py:57-296).
The engine does not bake a grammar in: it takes a table that maps each Symbol
value to an engine opcode (``ppi_symbol`` in include/ppi_engine.h).  This module
restates the numbering rule so the table and a reference-compatible ``symbol``
header can be produced from the same BNF file:

* non-terminals are numbered from 0 in order of appearance: ``NT_INICIAL`` for
  the start rule, then one ``NT_<NAME>`` per rule head (read_grammar.py:66-100);
* terminals are numbered from ``TERMINAL_MIN = 10000`` in order of first
  appearance; ``+ - * / < <= > >= = !=`` are spelled ``T_ADD T_SUB T_MULT T_DIV
  T_LESS T_LESSEQUAL T_GREATER T_GREATEREQUAL T_EQUAL T_NOTEQUAL``, every other
  terminal is ``T_`` + its upper-cased text (read_grammar.py:19-41,128-196);
* a terminal whose text contains ``attr`` is an attribute: the first one adds
  ``T_ATTRIBUTE`` (numbered right after the last ordinary terminal) and each
  ``attrK`` is ``T_ATTRK = ATTRIBUTE_MIN + K`` with ``ATTRIBUTE_MIN = 20000``
  (read_grammar.py:146-152,275-276); PPI's decoder rewrites ``T_ATTRK`` into
  ``T_ATTRIBUTE`` with ephemeral ``K`` (ppi.cc:167-171);
* ``NOT_USING_T_CONST`` is defined when the grammar has no ``const``
  (read_grammar.py:290-293);
* ``MAX_QUANT_SIMBOLOS_POR_REGRA`` is the longest alternative (read_grammar.py:224).
"""
from __future__ import annotations

import dataclasses
import sys
from typing import Dict, List, Tuple

TERMINAL_MIN = 10000
ATTRIBUTE_MIN = 20000

_OPERATOR_SPELLING = {
    ">": "GREATER", ">=": "GREATEREQUAL", "<": "LESS", "<=": "LESSEQUAL", "=": "EQUAL",
    "!=": "NOTEQUAL", "+": "ADD", "-": "SUB", "*": "MULT", "/": "DIV",
}

# Engine opcodes (include/ppi_engine.h, enum ppi_op).
OPS = {
    "UNKNOWN": 0, "CONST": 1, "ATTRIBUTE": 2, "VALUE": 3,
    "IF_THEN_ELSE": 10,
    "AND": 11, "OR": 12, "XOR": 13, "GREATER": 14, "GREATEREQUAL": 15, "LESS": 16, "LESSEQUAL": 17,
    "EQUAL": 18, "NOTEQUAL": 19, "ADD": 20, "SUB": 21, "MULT": 22, "MEAN": 23, "MAX": 24, "MIN": 25,
    "MOD": 26, "POW": 27, "DIV": 28,
    "NOT": 40, "ABS": 41, "POW2": 42, "POW3": 43, "POW4": 44, "POW5": 45, "NEG": 46, "ROUND": 47,
    "CEIL": 48, "FLOOR": 49, "EXP": 50, "EXP10": 51, "EXP2": 52, "SIN": 53, "COS": 54, "STEP": 55,
    "SIGN": 56, "LOGISTIC": 57, "GAMMA": 58, "GAUSS": 59, "SQRT": 60, "LOG": 61, "LOG10": 62,
    "LOG2": 63, "TAN": 64,
}

# Named constants pushed by interpreter_core:152-190 (float literals of the symbol header).
NAMED_CONSTANTS = {
    "PI": 3.14159274101257, "PI_2": 1.57079637050629, "PI_4": 0.78539818525314,
    "1_PI": 0.31830987334251, "2_PI": 0.63661974668503, "2_SQRTPI": 1.12837922573090,
    "SQRT2": 1.41421353816986, "SQRT1_2": 0.70710676908493, "E": 2.71828174591064,
    "LOG2E": 1.44269502162933, "LOG10E": 0.43429449200630, "LN2": 0.69314718246460,
    "LN10": 2.30258512496948,
}


def terminal_op(name: str) -> Tuple[int, float] | None:
    """Engine opcode (and constant value) for a terminal enum name such as ``T_ADD``.

    Integer pushes ``T_0..T_100`` and ``T_N1..T_N100`` push their value
    (interpreter_core:191-793).  Returns None for names the interpreter does not
    know, which the reference evaluates as NaN (sequential.cc:160-162).
    """
    if not name.startswith("T_"):
        return None
    body = name[2:]
    if body in OPS and body not in ("UNKNOWN", "VALUE"):
        return OPS[body], 0.0
    if body in NAMED_CONSTANTS:
        return OPS["VALUE"], NAMED_CONSTANTS[body]
    neg = body.startswith("N") and body[1:].isdigit()
    digits = body[1:] if neg else body
    if digits.isdigit() and 0 <= int(digits) <= 100 and not (neg and int(digits) == 0):
        return OPS["VALUE"], float(-int(digits) if neg else int(digits))
    return None


@dataclasses.dataclass
class Grammar:
    nonterminals: List[str]          # enum order, NT_ names
    terminals: List[str]             # enum order from TERMINAL_MIN, T_ names
    attributes: List[Tuple[str, int]]  # (T_ATTRK, K) in order of appearance
    max_rule: int
    rules: Dict[str, List[List[str]]]  # head -> alternatives (raw tokens)

    def values(self) -> Dict[str, int]:
        out: Dict[str, int] = {}
        for k, n in enumerate(self.nonterminals):
            out[n] = k
        nxt = len(self.nonterminals)
        for k, t in enumerate(self.terminals):
            out[t] = TERMINAL_MIN + k
            nxt = TERMINAL_MIN + k + 1
        if self.attributes:
            out["T_ATTRIBUTE"] = nxt
            for name, k in self.attributes:
                out[name] = ATTRIBUTE_MIN + k
        return out

    @property
    def uses_const(self) -> bool:
        return "T_CONST" in self.terminals

    def engine_table(self) -> List[Tuple[int, int, float]]:
        """(symbol value, engine opcode, constant value) for every evaluable terminal."""
        vals = self.values()
        table = []
        for t in self.terminals:
            m = terminal_op(t)
            if m is not None:
                table.append((vals[t], m[0], m[1]))
        if self.attributes:
            table.append((vals["T_ATTRIBUTE"], OPS["ATTRIBUTE"], 0.0))
        return table

    def rule_tables(self) -> Tuple[List[int], List[List[List[int]]]]:
        """The decoder's grammar (the ``grammar`` header's ``gramatica[]`` / ``tamanhos[]``,
        read_grammar.py:160-230): per non-terminal in enum order, its alternatives in BNF order,
        each a list of Symbol values (an attribute keeps its ``T_ATTRK`` value; PPI's decoder
        rewrites it, ppi.cc:167-171).  Returns (non-terminal values, alternatives)."""
        vals = self.values()
        nts, alts_all = [], []
        for name in self.nonterminals:
            head = name[3:]
            nts.append(vals[name])
            alts = []
            for alt in self.rules[head]:
                syms = []
                for t in alt:
                    if t.startswith("<") and t.endswith(">"):
                        syms.append(vals["NT_" + t[1:-1].upper()])
                    else:
                        syms.append(vals[_terminal_name(t)])
                alts.append(syms)
            alts_all.append(alts)
        return nts, alts_all

    def symbol_header(self) -> str:
        """C header equivalent to the reference's generated ``symbol`` file."""
        parts = list(self.nonterminals)
        for k, t in enumerate(self.terminals):
            parts.append(f"{t} = TERMINAL_MIN" if k == 0 else t)
        if self.attributes:
            parts.append("T_ATTRIBUTE")
            parts.extend(f"{n} = ATTRIBUTE_MIN + {k}" for n, k in self.attributes)
        head = ["#ifndef __SYMBOL_H", "#define __SYMBOL_H", "",
                f"#define TERMINAL_MIN {TERMINAL_MIN}", f"#define ATTRIBUTE_MIN {ATTRIBUTE_MIN}", "",
                "typedef enum { " + ", ".join(parts) + " } Symbol;"]
        if not self.uses_const:
            head += ["", "#define NOT_USING_T_CONST 1"]
        head += ["", "#endif", ""]
        return "\n".join(head)


def _terminal_name(tok: str) -> str:
    return "T_" + (_OPERATOR_SPELLING.get(tok) or tok.upper())


def parse_bnf(text: str) -> Grammar:
    """Parse PPI's BNF format (``S = <start> ...`` then ``P = <a> ::= x | y ...``)."""
    toks: List[str] = []
    for line in text.splitlines():
        if line.lstrip().startswith("#"):
            continue
        toks += line.split()
    nts: List[str] = []
    terms: List[str] = []
    attrs: List[Tuple[str, int]] = []
    rules: Dict[str, List[List[str]]] = {}
    max_rule = 0

    def see_terminal(tok: str) -> None:
        name = _terminal_name(tok)
        if tok not in _OPERATOR_SPELLING and "ATTR" in tok.upper():
            if name not in [a for a, _ in attrs]:
                attrs.append((name, int(tok[4:])))
        elif name not in terms:
            terms.append(name)

    if "S" not in toks or "P" not in toks:
        raise ValueError("grammar needs an 'S =' start rule and a 'P =' production section")
    i = toks.index("S")
    p = toks.index("P")
    start = [t for t in toks[i + 1:p] if t != "="]
    nts.append("NT_INICIAL")
    max_rule = max(max_rule, len(start))
    rules["INICIAL"] = [start]
    j = p + 1
    heads = [k for k in range(j, len(toks)) if toks[k] == "::="]
    for h, k in enumerate(heads):
        name = toks[k - 1][1:-1].upper()
        if "NT_" + name not in nts:
            nts.append("NT_" + name)
        end = heads[h + 1] - 1 if h + 1 < len(heads) else len(toks)
        body = toks[k + 1:end]
        alts: List[List[str]] = [[]]
        for t in body:
            if t == "|":
                alts.append([])
            else:
                alts[-1].append(t)
        rules[name] = alts
        for alt in alts:
            max_rule = max(max_rule, len(alt))
            for t in alt:
                if not (t.startswith("<") and t.endswith(">")):
                    see_terminal(t)
    return Grammar(nts, terms, attrs, max_rule, rules)


def main(argv: List[str]) -> int:
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("-g", "--grammar", required=True)
    ap.add_argument("-o", "--output-dir", required=True)
    a = ap.parse_args(argv)
    import os
    g = parse_bnf(open(a.grammar).read())
    os.makedirs(a.output_dir, exist_ok=True)
    with open(os.path.join(a.output_dir, "symbol"), "w") as f:
        f.write(g.symbol_header())
    with open(os.path.join(a.output_dir, "ppi_symbol_table.inc"), "w") as f:
        f.write("/* generated from %s: {symbol, engine opcode, value} */\n" % os.path.basename(a.grammar))
        for s, op, v in g.engine_table():
            f.write("{%d, %d, %.9ef},\n" % (s, op, v))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
