"""Test vectors for any grammar: the counterpart of `alpacc test generate` and
`alpacc test compare` (app/Main.hs:342-343).

generate writes one request batch in the wire format (docs/wire-protocols.md)
for a build mode:
  --exhaustive K   every string of length K over the lexer DFA's alphabet
                   (listProducts, src/Alpacc/Util.hs:23-24; Test/Lexer.hs:
                   404-407), or over the non-ignore terminal ids in parser
                   mode;
  --long N         inputs of N bytes from random walks of the token DFA that
                   enter each token by a producing transition and keep an
                   accepting state reachable, converging on one past a
                   16-byte token budget (Test/Lexer.hs:340-387);
  --derivations C  C random derivations of the grammar (Test/LexerParser.hs:
                   445-479): terminal-id strings in parser mode, else text
                   with a sample string per terminal (ignore text between
                   them when the grammar has an ignore terminal), half of
                   them with one token or byte changed.
compare checks two response streams for the same batch record by record
(one from the reference's generated program, one from alpacc-hip, say) and
names the first records that differ.

  python -m alpacc_amd.testvec generate GRAMMAR --mode both --long 4096 --count 8 -o req.bin
  python -m alpacc_amd.testvec compare GRAMMAR --mode both expected.bin actual.bin

Expected outputs come from running a backend on the batch; this module only
builds inputs and compares outputs.
"""

from __future__ import annotations

import argparse
import itertools
import os
import random
import struct
import sys
from collections import deque
from typing import Dict, List, Optional

from alpacc_amd import wire
from alpacc_amd.gen import tables

HERE = os.path.dirname(os.path.abspath(__file__))
GRAMMARS = os.path.join(os.path.dirname(HERE), "grammars")
TOKEN_BUDGET = 16


def load(grammar: str) -> tables.Grammar:
    path = grammar if os.path.exists(grammar) else os.path.join(GRAMMARS, grammar + ".alp")
    return tables.load(path)


def layout(g: tables.Grammar, mode: str) -> wire.Layout:
    if mode == "lexer":
        return wire.Layout("lexer", g.terminal_bytes_lex, g.index_bytes, g.length_bytes)
    if mode == "parser":
        return wire.Layout("parser", g.terminal_bytes_both, g.index_bytes, 0, g.parser.production_bytes)
    return wire.Layout("both", g.terminal_bytes_both, g.index_bytes, g.length_bytes, g.parser.production_bytes)


# ---------------------------------------------------------------- the DFA

def alphabet(g: tables.Grammar) -> List[int]:
    """Bytes with a transition out of some state."""
    dfa = g.lex_both.dfa
    return sorted({b for s in range(1, dfa.n + 1) for b in range(256) if dfa.delta[s][b]})


def _coreachable(dfa) -> List[int]:
    """Steps from each state to the nearest accepting state (-1: none)."""
    rev: Dict[int, set] = {s: set() for s in range(dfa.n + 1)}
    for s in range(1, dfa.n + 1):
        for b in range(256):
            t = dfa.delta[s][b]
            if t:
                rev[t].add(s)
    dist = [-1] * (dfa.n + 1)
    q = deque()
    for s in range(1, dfa.n + 1):
        if dfa.accepting(s):
            dist[s] = 0
            q.append(s)
    while q:
        t = q.popleft()
        for s in rev[t]:
            if dist[s] < 0:
                dist[s] = dist[t] + 1
                q.append(s)
    return dist


def dfa_walk(g: tables.Grammar, n: int, rng: random.Random) -> bytes:
    """n bytes (a few more to end in an accepting state) walked through the
    token DFA: every step keeps an accepting state reachable; past the token
    budget the walk takes a producing transition (a new token) when one
    exists, else the step nearest to acceptance."""
    dfa = g.lex_both.dfa
    dist = _coreachable(dfa)
    out = bytearray()
    s, tok_len = 1, 0
    while len(out) < n or not dfa.accepting(s):
        moves = [b for b in range(256) if dfa.delta[s][b] and dist[dfa.delta[s][b]] >= 0]
        if not moves:
            break
        if len(out) >= n or tok_len >= TOKEN_BUDGET:
            prod = [b for b in moves if dfa.produce[s][b]] if dfa.accepting(s) else []
            if prod and len(out) < n:
                moves = prod
            else:
                best = min(dist[dfa.delta[s][b]] for b in moves)
                moves = [b for b in moves if dist[dfa.delta[s][b]] == best]
        b = rng.choice(moves)
        tok_len = 1 if dfa.produce[s][b] else tok_len + 1
        out.append(b)
        s = dfa.delta[s][b]
    return bytes(out)


def terminal_texts(g: tables.Grammar) -> Dict[int, bytes]:
    """A shortest string the lexer reads as each terminal (combined ids)."""
    dfa = g.lex_both.dfa
    seen = {1: b""}
    q = deque([1])
    out: Dict[int, bytes] = {}
    while q:
        s = q.popleft()
        if dfa.accepting(s):
            t = g.lex_both.term_id[dfa.token[s]]
            out.setdefault(t, seen[s])
        for b in range(256):
            t = dfa.delta[s][b]
            if t and t not in seen and not dfa.produce[s][b]:
                seen[t] = seen[s] + bytes([b])
                q.append(t)
    return out


# ---------------------------------------------------------------- the CFG

def _heights(g: tables.Grammar) -> Dict[int, int]:
    """The least height of a derivation tree from each nonterminal to a
    terminal string (a fixpoint over the productions)."""
    nterm = g.parser.end_t + 1
    h: Dict[int, int] = {}
    changed = True
    while changed:
        changed = False
        for lhs, rhs in g.productions:
            if all(x < nterm or x in h for x in rhs):
                v = 1 + max([h[x] for x in rhs if x >= nterm], default=0)
                if v < h.get(lhs, 1 << 30):
                    h[lhs] = v
                    changed = True
    return h


def derivation(g: tables.Grammar, rng: random.Random, budget: int = 400) -> Optional[List[int]]:
    """Terminal ids of a random leftmost derivation from the start symbol
    (None when it outgrows its budget).  Past the budget every nonterminal
    takes an alternative of least height (then fewest symbols), so each
    expansion lowers the height left and the derivation ends."""
    prods = g.productions
    nterm = g.parser.end_t + 1
    height = _heights(g)
    by_lhs: Dict[int, list] = {}
    for lhs, rhs in prods:
        by_lhs.setdefault(lhs, []).append(rhs)

    def rank(rhs):
        return (max([height.get(x, 1 << 30) for x in rhs if x >= nterm], default=0), len(rhs))

    form, out = [prods[-1][1][1]], []
    while form:
        s = form.pop(0)
        if s < nterm:
            out.append(s)
            continue
        alts = by_lhs[s]
        budget -= 1
        if budget < 0:
            alts = [min(alts, key=rank)]
        form = list(rng.choice(alts)) + form
        if len(form) > 4 * 400:
            return None
    return out


# ---------------------------------------------------------------- batches

def generate(g: tables.Grammar, mode: str, exhaustive: int = 0, long: int = 0, derivations: int = 0,
             count: int = 1, seed: int = 0) -> List[bytes]:
    rng = random.Random(seed)
    if exhaustive:
        if mode == "parser":
            alpha = sorted(set(g.lex_both.term_id.values()) - {g.lex_both.ignore_id})
        else:
            alpha = alphabet(g)
        return [bytes(t) for t in itertools.product(alpha, repeat=exhaustive)]
    if long:
        items = [dfa_walk(g, long, rng) for _ in range(count)]
        if mode == "parser":  # the walks' kept tokens
            return [bytes(toks) for toks in (_lex_ids(g, x) for x in items) if toks is not None]
        return items
    out = []
    texts = terminal_texts(g)
    ign = g.lex_both.ignore_id
    sep = texts.get(ign, b"") if ign is not None else b""
    misses = 0
    while len(out) < derivations:
        toks = derivation(g, rng)
        if toks is None:
            misses += 1
            if misses > 100 * derivations + 1000:
                raise ValueError(f"{g.name}: random derivations keep outgrowing their budget")
            continue
        if mode == "parser":
            item = bytes(toks)
            if len(out) % 2 and item:
                b = bytearray(item)
                b[rng.randrange(len(b))] = rng.choice(sorted(set(g.lex_both.term_id.values())))
                item = bytes(b)
        else:
            item = sep.join(texts.get(t, b"") for t in toks)
            if len(out) % 2 and item:
                b = bytearray(item)
                b[rng.randrange(len(b))] = rng.choice(alphabet(g))
                item = bytes(b)
        out.append(item)
    return out


def _lex_ids(g: tables.Grammar, text: bytes) -> Optional[List[int]]:
    """Combined-mode terminal ids of text (kept tokens), None if rejected."""
    from alpacc_amd.gen import lexgen
    toks = lexgen.tokenize(g.lex_both.dfa, text)
    if toks is None:
        return None
    ids = [g.lex_both.term_id[t] for t, _, _ in toks]
    return [t for t in ids if t != g.lex_both.ignore_id]


def compare(expected: bytes, actual: bytes, lay: wire.Layout, limit: int = 5) -> List[str]:
    """Differences between two response streams (headers or records)."""
    msgs = []
    ne, na = struct.unpack_from("<Q", expected, 0)[0], struct.unpack_from("<Q", actual, 0)[0] if len(actual) >= 8 else -1
    if ne != na:
        return [f"record counts differ: expected {ne}, actual {na}"]
    oe = oa = 8
    for i in range(ne):
        re_, oe2 = wire.decode_record(expected, oe, lay)
        try:
            ra, oa2 = wire.decode_record(actual, oa, lay)
        except Exception as exc:  # a malformed record ends the comparison
            msgs.append(f"record {i}: actual undecodable ({exc})")
            break
        if re_.raw != ra.raw:
            what = "verdict" if re_.valid != ra.valid else "fields"
            msgs.append(f"record {i}: {what} differ (expected {len(re_.raw)} B, actual {len(ra.raw)} B)")
            if len(msgs) >= limit:
                break
        oe, oa = oe2, oa2
    return msgs


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m alpacc_amd.testvec")
    sub = ap.add_subparsers(dest="cmd", required=True)
    gp = sub.add_parser("generate")
    gp.add_argument("grammar")
    gp.add_argument("--mode", choices=["lexer", "parser", "both"], default="both")
    kind = gp.add_mutually_exclusive_group(required=True)
    kind.add_argument("--exhaustive", type=int, metavar="K")
    kind.add_argument("--long", type=int, metavar="N")
    kind.add_argument("--derivations", type=int, metavar="C")
    gp.add_argument("--count", type=int, default=1, help="inputs of --long")
    gp.add_argument("--seed", type=int, default=0)
    gp.add_argument("-o", "--output", required=True)
    cp = sub.add_parser("compare")
    cp.add_argument("grammar")
    cp.add_argument("--mode", choices=["lexer", "parser", "both"], default="both")
    cp.add_argument("expected")
    cp.add_argument("actual")
    a = ap.parse_args(argv)
    g = load(a.grammar)
    lay = layout(g, a.mode)
    if a.cmd == "generate":
        items = generate(g, a.mode, a.exhaustive or 0, a.long or 0, a.derivations or 0, a.count, a.seed)
        with open(a.output, "wb") as f:
            f.write(wire.encode_batch(items, lay))
        print(f"{len(items)} requests, {sum(map(len, items))} payload bytes -> {a.output}")
        return 0
    exp = open(a.expected, "rb").read()
    act = open(a.actual, "rb").read()
    msgs = compare(exp, act, lay)
    for m in msgs:
        print(m)
    print("match" if not msgs else "MISMATCH")
    return 0 if not msgs else 1


if __name__ == "__main__":
    sys.exit(main())
