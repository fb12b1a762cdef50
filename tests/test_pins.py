"""Known answers that pin the oracle's numbering, its lexer and the library's
names, independently of the table generator that both share.

* The reference's numbering rules (src/Alpacc/Encode.hs:118-158,
  src/Alpacc/Generator/Analyzer.hs:113-126, src/Alpacc/Lexer/DFA.hs:312-313,
  src/Alpacc/Grammar.hs:191-211, 339-361): lexer-only terminal ids follow
  the sorted terminal set, combined ids the grammar's order, so json.alp
  numbers `number` 0 and `string` 1 in lexer mode and the other way round in
  combined mode; hard-coded tables for arithmetic, json and sexp.
* Records: a JSON lexer-only record, its combined counterpart, and the
  arithmetic input `3-5`, which lexes as `num num` (num = /-?[0-9]+/ and
  greedy lexing, grammars/arithmetic.alp:9) and so does not parse.
* The lexer against an implementation that shares nothing with the
  generator's DFA: every terminal's regex is translated into a pattern of
  Python's `regex` module, whose partial matching says whether a string can
  still grow into a token.  A token runs while its text is a prefix of some
  terminal's language (the DFA continues whenever it can, DFA.hs:261-299) and
  is then the lowest-priority terminal (literals first, then declared rules
  in file order, CFG.hs:138-162) whose language holds the text, else the
  input is rejected.
* A numbering planted wrong in a copy of the analysed grammar fails the
  known-answer check (the check can tell the two numberings apart).
"""

import copy
import functools
import itertools
import os

import pytest
import regex

import pyoracle
from alpacc_amd import wire
from alpacc_amd.gen import tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@functools.lru_cache(maxsize=None)
def load(name):
    return tables.load(os.path.join(ROOT, "grammars", name + ".alp"))


# ----------------------------------------------------------- numbering tables

# terminal ids, combined / parser numbering: name -> id (dead last)
BOTH_IDS = {
    "arithmetic": ["num", "var", "(", ")", "*", "+", "-", "/", "ignore", "<dead>"],
    "json": ["string", "number", ",", ":", "[", "]", "false", "null", "true", "{", "}", "ignore", "<dead>"],
    "sexp": ["atom", "(", ")", "ignore", "<dead>"],
}
LEX_IDS = {
    "arithmetic": BOTH_IDS["arithmetic"],
    "json": ["number", "string", ",", ":", "[", "]", "false", "null", "true", "{", "}", "ignore", "<dead>"],
    "sexp": BOTH_IDS["sexp"],
}
# productions: (name as the Futhark target names it, arity)
PRODUCTIONS = {
    "arithmetic": [("E0_0", 2), ("Add", 3), ("Sub", 3), ("Empty", 0), ("E1_4", 2), ("Mul", 3), ("Div", 3),
                   ("Empty", 0), ("Parentheses", 3), ("Num", 1), ("Var", 1),
                   ("num_11", 0), ("var_12", 0), ("ignore_13", 0)] + [(f"Literal_{i}", 0) for i in range(14, 20)]
                  + [("start_20", 1)],
    "json": [("Object_production", 1), ("Array_production", 1), ("String", 1), ("Number", 1), ("Null", 1),
             ("True", 1), ("False", 1), ("Object_head", 3), ("Object_tail", 0), ("Object_tail", 2),
             ("Object_elem", 0), ("Object_elem", 3), ("Object_key", 3), ("Array_head", 3), ("Array_tail", 0),
             ("Array_tail", 2), ("Array_elem", 0), ("Array_elem", 3),
             ("string_18", 0), ("number_19", 0), ("ignore_20", 0)] + [(f"Literal_{i}", 0) for i in range(21, 30)]
            + [("start_30", 1)],
    "sexp": [("E_0", 1), ("E_1", 3), ("E0_2", 0), ("E0_3", 2), ("ignore_4", 0), ("atom_5", 0), ("Literal_6", 0),
             ("Literal_7", 0), ("start_8", 1)],
}
WIDTHS = {  # terminal_t, production_t, index_t, length_t
    "arithmetic": (1, 1, 8, 8), "json": (1, 1, 8, 2), "sexp": (1, 1, 8, 2),
}


def term_text(t):
    return t[1]


def numbering_errors(g, name):
    """Every way the analysed grammar differs from the tables above."""
    errs = []
    for lt, exp, what in ((g.lex_both, BOTH_IDS[name], "combined"), (g.lex_only, LEX_IDS[name], "lexer")):
        got = {term_text(t): i for t, i in lt.term_id.items()}
        for i, nm in enumerate(exp[:-1]):
            if got.get(nm) != i:
                errs.append(f"{what}: {nm} is {got.get(nm)}, not {i}")
        if lt.dead_id != len(exp) - 1:
            errs.append(f"{what}: dead terminal {lt.dead_id}")
    p = g.parser
    got = list(zip(p.prod_names, p.arity))
    if got != PRODUCTIONS[name]:
        errs.append(f"productions {got}")
    if (g.terminal_bytes_both, p.production_bytes, g.index_bytes, g.length_bytes) != WIDTHS[name]:
        errs.append("widths")
    return errs


@pytest.mark.parametrize("name", sorted(BOTH_IDS))
def test_numbering_known_answers(name):
    assert numbering_errors(load(name), name) == []


def test_planted_grammar_order_numbering_is_caught():
    """A generator numbering lexer mode in grammar order (json: string 0,
    number 1) fails the check above."""
    g = copy.deepcopy(load("json"))
    ids = g.lex_only.term_id
    s, n = ("T", "string"), ("T", "number")
    ids[s], ids[n] = ids[n], ids[s]
    assert any("lexer" in e for e in numbering_errors(g, "json"))


# ------------------------------------------------------------------- records

def rec(grammar, mode, payload):
    g = load(grammar)
    lay = (wire.Layout("lexer", g.terminal_bytes_lex, g.index_bytes, g.length_bytes) if mode == "lexer" else
           wire.Layout("both", g.terminal_bytes_both, g.index_bytes, g.length_bytes, g.parser.production_bytes))
    raw = pyoracle.record(grammar, mode, payload)
    r, off = wire.decode_record(raw, 0, lay)
    assert off == len(raw)
    return r


def lexemes(r):
    return list(zip(r.terminals.tolist(), r.starts.tolist(), r.lengths.tolist()))


def test_json_lexer_only_record():
    r = rec("json", "lexer", b'"a" 1')
    assert r.valid and lexemes(r) == [(1, 0, 3), (0, 4, 1)]


def test_json_combined_record():
    r = rec("json", "both", b'["a",1]')
    assert r.valid and r.terminals.tolist() == [4, 0, 2, 1, 5]
    # J -> A; A -> [ EL0 ]; EL0 -> J EL1; J -> String; EL1 -> , J EL1; J -> Number; EL1 -> (empty)
    assert r.nodes.tolist() == [1, 13, 15, 2, 17, 3, 16]
    assert r.parents.tolist() == [0, 0, 1, 2, 2, 4, 4]
    assert r.token_parents.tolist() == [1, 3, 4, 5, 1]


def test_arithmetic_minus_digit_is_one_number():
    r = rec("arithmetic", "lexer", b"3-5")
    assert r.valid and lexemes(r) == [(0, 0, 1), (0, 1, 2)]
    assert not rec("arithmetic", "both", b"3-5").valid
    r = rec("arithmetic", "both", b"3- 5")
    assert r.valid and r.terminals.tolist() == [0, 6, 0]


# ------------------------------------------------- lexer from the regexes alone

def to_pattern(rx) -> bytes:
    """The parsed regex AST (grammar.py) as a pattern of the regex module."""
    k = rx[0]
    if k == "eps":
        return b""
    if k == "lit":
        return regex.escape(rx[1])
    if k == "range":
        return b"(?:" + b"|".join(regex.escape(x) for x in rx[1]) + b")"
    if k == "star":
        return b"(?:" + to_pattern(rx[1]) + b")*"
    if k == "alt":
        return b"(?:" + to_pattern(rx[1]) + b"|" + to_pattern(rx[2]) + b")"
    return to_pattern(rx[1]) + to_pattern(rx[2])


def regex_lexer(g):
    """(priority-ordered (terminal, compiled pattern)) from the grammar text."""
    from alpacc_amd.gen import grammar as G
    cfg = g.cfg
    declared = [r.term for r in cfg.t_rules]
    lits = sorted({s[1] for r in cfg.nt_rules for s in r.rhs if s[0] == "T" and s[1] not in declared},
                  key=lambda t: (0 if t[0] == "T" else 1, t[1]))
    rules = [(t, G.literal_regex(t[1])) for t in lits] + [(r.term, r.regex) for r in cfg.t_rules]
    return [(t, regex.compile(to_pattern(rx))) for t, rx in rules]


def regex_lex(g, pats, data: bytes, lexer_mode=True):
    lt = g.lex_only if lexer_mode else g.lex_both
    out, pos, n = [], 0, len(data)
    if n == 0:
        return []
    while pos < n:
        end = pos
        while end < n and any(p.fullmatch(data[pos:end + 1], partial=True) for _, p in pats):
            end += 1
        text = data[pos:end]
        match = next((t for t, p in pats if text and p.fullmatch(text)), None)
        if match is None:
            return None
        tid = lt.term_id[match]
        if tid != lt.ignore_id:
            if end - pos > (1 << (8 * g.length_bytes)) - 1:
                return None
            out.append((tid, pos, end - pos))
        pos = end
    return out


ALPHABETS = {"arithmetic": b"1-a(+ )*/", "json": b'"a1 [,]{:}-.tn', "sexp": b"a1 ()\n",
             "paper_grammar": b"a+[] ", "numfmt": b"0x1.e/,b", "lexer": b'1 "\\a'}


@pytest.mark.parametrize("name", sorted(ALPHABETS))
def test_lexer_matches_regex_semantics(name):
    g = load(name)
    pats = regex_lexer(g)
    alpha = ALPHABETS[name]
    maxn = 4 if len(alpha) > 8 else 5
    for n in range(0, maxn + 1):
        for tup in itertools.product(alpha, repeat=n):
            data = bytes(tup)
            exp = regex_lex(g, pats, data)
            r = rec(name, "lexer", data) if g.has_parser else lexer_only_rec(name, data)
            assert r.valid == (exp is not None), data
            if exp is not None:
                assert lexemes(r) == exp, data


def lexer_only_rec(name, data):
    g = load(name)
    lay = wire.Layout("lexer", g.terminal_bytes_lex, g.index_bytes, g.length_bytes)
    raw = pyoracle.record(name, "lexer", data)
    r, off = wire.decode_record(raw, 0, lay)
    assert off == len(raw)
    return r


def test_lexer_only_grammar():
    """grammars/lexer.alp has terminal rules alone: a lexer in lexer mode
    (sorted numbering: number 0, string 1, ignore 2), and no parser."""
    g = load("lexer")
    assert not g.has_parser and g.modes() == ["lexer"]
    assert {term_text(t): i for t, i in g.lex_only.term_id.items()} == {"number": 0, "string": 1, "ignore": 2}
    r = lexer_only_rec("lexer", b'12 "a\\"b" 7')
    assert r.valid and lexemes(r) == [(0, 0, 2), (1, 3, 6), (0, 10, 1)]
    assert not lexer_only_rec("lexer", b'"open').valid
