"""Per-grammar analysis: numbering, integer widths, lexer and parser tables.

Numbering rules (all wire-visible) restate the reference analyzer:

* grammar terminals: declared terminals in file order, then terminals used in
  productions in ``Set`` order (named before literals, then by text), without
  duplicates -- src/Alpacc/CFG.hs:112-136;
* combined / parser-only terminal ids: non-ignore terminals in grammar order,
  then ``ignore``, the dead terminal, the start sentinel, the end sentinel;
  nonterminal symbol ids follow -- src/Alpacc/Encode.hs:118-142;
* lexer-only terminal ids: terminals sorted, ``ignore`` and the dead terminal
  last -- src/Alpacc/Generator/Analyzer.hs:113-126, Encode.hs:144-158;
* productions: user productions in file order (one per alternative), one
  terminal production per grammar terminal, then ``Start -> |- S -|`` --
  src/Alpacc/Grammar.hs:339-361, 191-211; arity = number of RHS symbols
  (Analyzer.hs:151-156);
* widths: terminal_t holds the largest terminal id, production_t the number
  of productions, index_t is i64 (or i32 with index32), length_t the grammar's
  ``length`` or the unsigned index type -- Encode.hs:217-247,
  Generator/C/Generator.hs:36-53.
"""

from __future__ import annotations

import dataclasses
import hashlib
import json
from typing import Dict, List, Optional, Tuple

from . import grammar as G
from . import lexgen
from . import llp

FNV_OFFSET = 14695981039346656037
FNV_PRIME = 1099511628211
M64 = (1 << 64) - 1


def uint_bytes(v: int) -> int:
    """Smallest unsigned width (bytes) holding v (Types.hs:53-58)."""
    for nb in (1, 2, 4, 8):
        if v < (1 << (8 * nb)):
            return nb
    raise ValueError("value too large for u64")


def bit_size(x: int) -> int:
    """findBitSize (Lexer/Encode.hs:42-45): bits to hold x-1, at least 1."""
    return max(1, x - 1).bit_length()


def fnv1a(key) -> int:
    h = FNV_OFFSET
    for t in key:
        h = ((h ^ t) * FNV_PRIME) & M64
    return h


@dataclasses.dataclass
class LexerTables:
    """Everything both lexer paths need for one terminal numbering."""

    dfa: lexgen.TokenDFA
    endos: lexgen.Endos
    term_id: Dict[tuple, int]          # terminal key -> id in this numbering
    ignore_id: Optional[int]
    dead_id: int
    n_terminals: int                   # size of the terminal encoder
    # oracle packing (Lexer/Encode.hs:33-141)
    endo_bits: int
    term_bits: int
    state_bytes: int
    to_state: List[int]
    compose: List[int]
    accept: List[bool]
    identity: int

    def names(self) -> List[str]:
        """Name of every terminal id, the dead one last (nameTerminals,
        src/Alpacc/Generator/Analyzer.hs:104-110): a named terminal by its
        name, a literal as literal_<id>, the dead terminal as empty_<id>."""
        out = [""] * (self.dead_id + 1)
        for t, i in self.term_id.items():
            out[i] = t[1] if t[0] == "T" else f"literal_{i}"
        out[self.dead_id] = f"empty_{self.dead_id}"
        return out


@dataclasses.dataclass
class ParserTables:
    q: int
    k: int
    empty_t: int
    start_t: int
    end_t: int
    n_symbols: int
    bracket_bytes: int
    production_bytes: int
    n_productions: int
    arity: List[int]
    prod_terminal: List[Optional[int]]
    prod_names: List[str]
    keys: List[Tuple[int, ...]]                      # padded q+k keys, sorted
    values: List[Tuple[List[int], List[int]]]        # (encoded brackets, productions)
    stacks: List[int]
    productions: List[int]
    spans: List[Tuple[int, int, int, int]]          # per key: stack s,e, prod s,e
    hash_size: int
    hash_valid: List[bool]
    hash_keys: List[Tuple[int, ...]]
    hash_spans: List[Tuple[int, int, int, int]]
    max_iters: int
    max_brackets: int
    max_prods: int


@dataclasses.dataclass
class Grammar:
    name: str
    cfg: G.CFG
    terminals: List[tuple]             # grammar terminal order
    q: int
    k: int
    length_bits: Optional[int]
    index32: bool
    lex_both: LexerTables              # combined / parser numbering
    lex_only: LexerTables              # lexer-only numbering
    parser: Optional[ParserTables]     # None: terminal rules only (a lexer-only grammar)
    terminal_bytes_both: int
    terminal_bytes_lex: int
    productions: List[Tuple[int, Tuple[int, ...]]] = dataclasses.field(default_factory=list)
    llp_grammar: Optional[llp.LLPGrammar] = None

    @property
    def index_bytes(self) -> int:
        return 4 if self.index32 else 8

    @property
    def length_bytes(self) -> int:
        if self.length_bits is None:
            return self.index_bytes
        return self.length_bits // 8

    @property
    def has_parser(self) -> bool:
        return self.parser is not None

    def modes(self) -> List[str]:
        """Build modes this grammar has: every mode, or the lexer alone."""
        return ["lexer", "parser", "both"] if self.has_parser else ["lexer"]

    def digest(self) -> str:
        blob = json.dumps({
            "to_state": self.lex_both.to_state, "n": self.lex_both.dfa.n,
            "keys": self.parser.keys if self.parser else [], "values": self.parser.values if self.parser else [],
        }).encode()
        return hashlib.sha256(blob).hexdigest()[:16]


def term_order(t: tuple):
    """Haskell's derived ``Ord T``: named terminals (T) before literals (TLit),
    then by text."""
    return (0 if t[0] == "T" else 1, t[1])


def _grammar_terminals(cfg: G.CFG) -> List[tuple]:
    declared = [r.term for r in cfg.t_rules]
    used = sorted({sym[1] for r in cfg.nt_rules for sym in r.rhs if sym[0] == "T"}, key=term_order)
    out: List[tuple] = []
    for t in declared + used:
        if t not in out:
            out.append(t)
    return out


def _lexer_rules(cfg: G.CFG):
    """Implicit literal rules (Set order) then declared rules; their index is
    the priority (CFG.hs:138-162)."""
    declared = [r.term for r in cfg.t_rules]
    implicit = sorted({sym[1] for r in cfg.nt_rules for sym in r.rhs
                       if sym[0] == "T" and sym[1] not in declared}, key=term_order)
    rules = []
    for t in implicit:
        if t[0] != "L":
            raise G.GrammarError(f"Can not create literal from: {t[1]}")
        rules.append((t, G.literal_regex(t[1])))
    for r in cfg.t_rules:
        rules.append((r.term, r.regex))
    for t, rx in rules:
        if G.produces_epsilon(rx):
            raise G.GrammarError(f"Error: {t} may not produce empty strings.")
    if not rules:
        raise G.GrammarError("CFG has no lexical rules.")
    order = {}
    for i, (t, _) in enumerate(rules):
        order[t] = i
    return rules, order


def _lexer_tables(dfa, endos, term_id, ignore_id, dead_id, n_terminals) -> LexerTables:
    N = endos.size
    eb = bit_size(N)
    tb = bit_size(n_terminals)
    total = eb + tb + 1
    state_bytes = uint_bytes((1 << total) - 1)

    def pack(e: int) -> int:
        f = endos.funcs[e]
        s = f[1]
        tok = dfa.token[s] if s else None
        t = term_id[tok] if tok is not None else dead_id
        p = 1 if endos.prods[e][1] else 0
        return e | (t << eb) | (p << (eb + tb))

    packed = [pack(e) for e in range(N)]
    to_state = [packed[endos.byte_endo[b]] for b in range(256)]
    compose = [packed[endos.compose[a][b]] for a in range(N) for b in range(N)]
    accept = []
    for e in range(N):
        s = endos.funcs[e][1]
        accept.append(bool(s) and dfa.token[s] is not None)
    return LexerTables(dfa, endos, term_id, ignore_id, dead_id, n_terminals, eb, tb,
                       state_bytes, to_state, compose, accept, packed[lexgen.IDENTITY_E])


def analyse(name: str, text: str, index32: bool = False, q: Optional[int] = None,
            k: Optional[int] = None) -> Grammar:
    """Analyse grammar text; ``q`` / ``k`` override the grammar's params."""
    cfg = G.parse_cfg(text)
    if q is not None:
        cfg.params.lookback = q
    if k is not None:
        cfg.params.lookahead = k
    q, k = cfg.params.lookback, cfg.params.lookahead
    terminals = _grammar_terminals(cfg)
    ignore = ("T", "ignore")

    # ---- lexer DFA (shared by both numberings) ----
    rules, order = _lexer_rules(cfg)
    dfa = lexgen.build_token_dfa(rules, order)
    endos = lexgen.endomorphisms(dfa)

    # ---- combined / parser numbering (Encode.hs:118-142) ----
    fts = [t for t in terminals if t != ignore]
    both_ids: Dict[tuple, int] = {}
    for t in fts:
        both_ids[t] = len(both_ids)
    ignore_both = None
    if ignore in terminals:
        ignore_both = len(both_ids)
        both_ids[ignore] = ignore_both
    dead_both = len(both_ids)
    start_t = dead_both + 1
    end_t = dead_both + 2
    n_term_syms = end_t + 1
    lex_both = _lexer_tables(dfa, endos, both_ids, ignore_both, dead_both, len(both_ids) + 1)

    # ---- lexer-only numbering (Encode.hs:144-158; terminals sorted) ----
    lex_terms = sorted(order.keys(), key=term_order)
    lex_ids: Dict[tuple, int] = {}
    for t in lex_terms:
        if t != ignore:
            lex_ids[t] = len(lex_ids)
    ignore_lex = None
    if ignore in lex_terms:
        ignore_lex = len(lex_ids)
        lex_ids[ignore] = ignore_lex
    dead_lex = len(lex_ids)
    lex_only = _lexer_tables(dfa, endos, lex_ids, ignore_lex, dead_lex, len(lex_ids) + 1)

    if not cfg.nt_rules:
        # a grammar of terminal rules alone builds a lexer only: the
        # reference's mkLexer needs no productions (Analyzer.hs:113-149),
        # while its parser and combined builds stop with "CFG has no
        # production rules." (CFG.hs:115)
        return Grammar(
            name=name, cfg=cfg, terminals=terminals, q=q, k=k, length_bits=cfg.params.length,
            index32=index32, lex_both=lex_both, lex_only=lex_only, parser=None,
            terminal_bytes_both=uint_bytes(end_t), terminal_bytes_lex=uint_bytes(dead_lex),
        )

    # ---- augmented grammar (Grammar.hs:191-381) ----
    user_nts: List[str] = []
    for r in cfg.nt_rules:
        if r.nt not in user_nts:
            user_nts.append(r.nt)
    # nonterminal symbols: one per grammar terminal, the user nonterminals, Start
    nt_sym: Dict[tuple, int] = {}
    for t in terminals:
        nt_sym[("T", t)] = n_term_syms + len(nt_sym)
    for nt in user_nts:
        nt_sym[("N", nt)] = n_term_syms + len(nt_sym)
    START = n_term_syms + len(nt_sym)
    n_symbols = START + 1
    for r in cfg.nt_rules:
        for sym in r.rhs:
            if sym[0] == "N" and ("N", sym[1]) not in nt_sym:
                raise G.GrammarError(f"undefined nonterminal {sym[1]}")
    prods: List[Tuple[int, Tuple[int, ...]]] = []
    names: List[str] = []
    arity: List[int] = []
    prod_terminal: List[Optional[int]] = []
    for i, r in enumerate(cfg.nt_rules):
        rhs = tuple(nt_sym[("T", s[1])] if s[0] == "T" else nt_sym[("N", s[1])] for s in r.rhs)
        prods.append((nt_sym[("N", r.nt)], rhs))
        names.append(r.name or f"{r.nt}_{i}")
        arity.append(len(rhs))
        prod_terminal.append(None)
    for t in terminals:
        i = len(prods)
        prods.append((nt_sym[("T", t)], (both_ids[t],)))
        names.append(f"{t[1]}_{i}" if t[0] == "T" else f"Literal_{i}")
        arity.append(0)
        prod_terminal.append(both_ids[t])
    start_nt = nt_sym[("N", cfg.nt_rules[0].nt)]
    prods.append((START, (start_t, start_nt, end_t)))
    names.append(f"start_{len(prods) - 1}")
    arity.append(1)
    prod_terminal.append(None)
    _check_productive(prods, list(nt_sym.values()) + [START], n_term_syms)

    nonterminals = list(nt_sym.values()) + [START]
    g = llp.LLPGrammar(n_term_syms, prods, nonterminals, START, q, k)
    table, starts = llp.llp_table(g)
    btable = llp.bracket_table(table, starts)

    bracket_bytes = uint_bytes(2 * (n_symbols - 1))
    top_bit = 8 * bracket_bytes - 1

    def enc(b):
        kind, s = b
        return s | (1 << top_bit) if kind == "L" else s

    def pad(x, y):
        x = tuple(x)[:q]
        y = tuple(y)[:k]
        return (dead_both,) * (q - len(x)) + x + y + (dead_both,) * (k - len(y))

    keyed = {}
    for (x, y), (br, ps) in btable.items():
        keyed[pad(x, y)] = ([enc(b) for b in br], list(ps))
    keys = sorted(keyed)
    values = [keyed[kk] for kk in keys]
    stacks: List[int] = []
    productions: List[int] = []
    spans = []
    for br, ps in values:
        s0 = len(stacks)
        stacks.extend(br)
        p0 = len(productions)
        productions.extend(ps)
        spans.append((s0, len(stacks), p0, len(productions)))

    # open addressing, FNV-1a, table size twice the key count (HashTable.hs:38-49)
    size = 2 * len(keys) if keys else 1
    hv = [False] * size
    hk = [tuple([0] * (q + k))] * size
    hs = [(-1, -1, -1, -1)] * size
    max_iters = 0
    for kk, sp in zip(keys, spans):
        h = fnv1a(kk) % size
        p = 0
        while hv[(h + p) % size]:
            p += 1
        hv[(h + p) % size] = True
        hk[(h + p) % size] = kk
        hs[(h + p) % size] = sp
        max_iters = max(max_iters, p + 1)

    parser = ParserTables(
        q=q, k=k, empty_t=dead_both, start_t=start_t, end_t=end_t, n_symbols=n_symbols,
        bracket_bytes=bracket_bytes, production_bytes=uint_bytes(len(prods)),
        n_productions=len(prods), arity=arity, prod_terminal=prod_terminal, prod_names=names,
        keys=keys, values=values, stacks=stacks, productions=productions, spans=spans,
        hash_size=size, hash_valid=hv, hash_keys=hk, hash_spans=hs, max_iters=max_iters,
        max_brackets=max([s[1] - s[0] for s in spans] + [0]),
        max_prods=max([s[3] - s[2] for s in spans] + [0]),
    )
    return Grammar(
        name=name, cfg=cfg, terminals=terminals, q=q, k=k, length_bits=cfg.params.length,
        index32=index32, lex_both=lex_both, lex_only=lex_only, parser=parser,
        terminal_bytes_both=uint_bytes(end_t), terminal_bytes_lex=uint_bytes(dead_lex),
        productions=prods, llp_grammar=g,
    )


def _check_productive(prods, nts, n_terms):
    """Grammar.hs:275-300: every nonterminal must derive a terminal string."""
    productive = set()
    changed = True
    while changed:
        changed = False
        for lhs, rhs in prods:
            if lhs not in productive and all(s < n_terms or s in productive for s in rhs):
                productive.add(lhs)
                changed = True
    bad = [nt for nt in nts if nt not in productive]
    if bad:
        raise G.GrammarError(f"nonproductive nonterminals: {bad}")


def load(path: str, index32: bool = False) -> Grammar:
    import os
    with open(path, encoding="utf-8") as f:
        text = f.read()
    name = os.path.splitext(os.path.basename(path))[0]
    return analyse(name, text, index32=index32)
