"""Random grammars for differential testing: the counterpart of `alpacc
random` (src/Alpacc/Random.hs:18-37, app/Main.hs:300-310, 540-548).

As there, a grammar has `num_terminals` terminals named a, b, c, ... whose
regexes are random expressions of size `num_terminals` over the first
`num_chars` lowercase letters (genDfaLexerSpec, src/Alpacc/Lexer/DFA.hs:
542-562; genRegEx, src/Alpacc/Lexer/RegularExpression.hs:235-259; a regex
that accepts the empty string is drawn again), and `num_nonterminals`
nonterminals A, B, ... with one production each plus `num_productions - n`
more on random left-hand sides; a right-hand side has 0..size/5 symbols,
the first a terminal (genGrammar, src/Alpacc/Grammar.hs:395-450).  A grammar
whose nonterminals are not all productive and reachable is drawn again.
With no productions asked for, the grammar has terminal rules only (a
lexer-only grammar).  The randomness is Python's seeded generator, so a
seed names a grammar.
"""

from __future__ import annotations

import random
from typing import List, Tuple


def _alpha(n: int, base: str) -> str:
    """intToAlpha: a, b, ..., z, aa, ab, ... (base 'a' or 'A')."""
    s = chr(ord(base) + n % 26)
    return _alpha(n // 26 - 1, base) + s if n >= 26 else s


def _regex(rng: random.Random, cs: List[str], size: int) -> str:
    """A random regex in .alp syntax (genRegEx)."""
    def rng_set():
        items = [rng.choice(cs)] + [rng.choice(cs) for _ in range(size)]
        return "[" + "".join(items) + "]"
    leaves = ["", "lit", "range"]
    kinds = leaves if size <= 1 else leaves + ["star", "cat", "alt"]
    k = rng.choice(kinds)
    if k == "":
        return "()"
    if k == "lit":
        return rng.choice(cs)
    if k == "range":
        return rng_set()
    if k == "star":
        return "(" + _regex(rng, cs, size // 2) + ")*"
    if k == "cat":
        return "(" + _regex(rng, cs, size // 2) + " " + _regex(rng, cs, size // 2) + ")"
    return "(" + _regex(rng, cs, size // 2) + "|" + _regex(rng, cs, size // 2) + ")"


def _nullable(rx: str) -> bool:
    from . import grammar as G
    parsed = G.parse_cfg(f"t = /{rx}/.\n").t_rules[0].regex
    return G.produces_epsilon(parsed)


def _closed(nts: List[str], prods: List[Tuple[str, List[str]]], start: str) -> bool:
    """Every nonterminal productive and reachable from the start."""
    productive = set()
    changed = True
    while changed:
        changed = False
        for lhs, rhs in prods:
            if lhs not in productive and all(s not in nts or s in productive for s in rhs):
                productive.add(lhs)
                changed = True
    reach = {start}
    changed = True
    while changed:
        changed = False
        for lhs, rhs in prods:
            if lhs in reach:
                for s in rhs:
                    if s in nts and s not in reach:
                        reach.add(s)
                        changed = True
    return productive == set(nts) and reach == set(nts)


def random_grammar(seed: int, num_chars: int = 3, num_terminals: int = 3, num_nonterminals: int = 2,
                   num_productions: int = 3, size: int = 30, q: int = None, k: int = None) -> str:
    """The .alp text of one random grammar (with a params block when q / k
    are given, as the reference's test scripts prepend one)."""
    rng = random.Random(seed)
    cs = [chr(ord("a") + i) for i in range(num_chars)]
    lines = []
    if q is not None or k is not None:
        lines += ["params {", f"  lookback = {q or 0}.", f"  lookahead = {1 if k is None else k}.", "}", ""]
    terms = []
    for i in range(num_terminals):
        while True:
            rx = _regex(rng, cs, num_terminals)
            if not _nullable(rx):
                break
        name = _alpha(i, "a")
        terms.append(name)
        lines.append(f"{name} = /{rx}/.")
    if num_nonterminals <= 0 or num_productions <= 0:
        return "\n".join(lines) + "\n"
    nts = [_alpha(i, "A") for i in range(max(1, num_nonterminals))]

    def rhs():
        n = rng.randint(0, size // 5)
        if n == 0:
            return []
        return [rng.choice(terms)] + [rng.choice(terms + nts) for _ in range(n - 1)]

    while True:
        start = rng.choice(nts)
        prods = [(nt, rhs()) for nt in nts]
        prods += [(rng.choice(nts), rhs()) for _ in range(num_productions - len(nts))]
        if _closed(nts, prods, start):
            break
    # the start symbol's production first: the grammar starts at the first rule
    prods.sort(key=lambda p: p[0] != start)
    lines.append("")
    lines += [f"{lhs} -> {' '.join(r)}." for lhs, r in prods]
    return "\n".join(lines) + "\n"
