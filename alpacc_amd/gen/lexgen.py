"""Lexer construction: regexes -> greedy tokenizing DFA -> byte endomorphisms.

Semantics follow the reference's lexer generator:

* each terminal's regex becomes a DFA; the DFAs are unioned by a subset
  construction; a union state's token is the accepting terminal of lowest
  *order* (src/Alpacc/Lexer/DFA.hs:322-395, orders from CFG.hs:138-162:
  production literals first in terminal order, then declared rules in file
  order);
* dead (useless) states are dropped, so a missing transition means reject
  (DFA.hs:196-213);
* lexing is greedy with no backtracking: from an accepting state ``q`` with
  no transition on byte ``b`` but with a transition from the initial state
  on ``b``, a *producing* transition ``q -b-> delta(init, b)`` ends the
  current token (DFA.hs:261-299).

The DFA built here is minimised with respect to everything observable (token,
accept, transitions and produce flags), so its state numbering is our own;
only the language-level behaviour has to match the reference.

The parallel form is the transformation semigroup of the DFA: every byte is an
endomorphism ``f_b`` of the state set (dead = 0) together with its produce
flags, composition is "apply the left one first" (ParallelLexing.hs:118-137,
DFAParallelLexer.hs:95-108), and a composed endomorphism's token / accept /
produce bits are those of the state it maps the initial state to
(ParallelLexing.hs:55-73).
"""

from __future__ import annotations

import dataclasses
from typing import Dict, FrozenSet, List, Optional, Sequence, Tuple

# --------------------------------------------------------------------------
# NFA / DFA over bytes
# --------------------------------------------------------------------------


class _NFA:
    def __init__(self):
        self.eps: List[List[int]] = []
        self.trans: List[Dict[int, List[int]]] = []

    def new(self) -> int:
        self.eps.append([])
        self.trans.append({})
        return len(self.eps) - 1

    def edge(self, a: int, byte: int, b: int):
        self.trans[a].setdefault(byte, []).append(b)

    def build(self, r, a: int, b: int):
        """Thompson construction of regex ``r`` between states a and b."""
        tag = r[0]
        if tag == "eps":
            self.eps[a].append(b)
        elif tag == "lit":
            self._bytes(a, r[1], b)
        elif tag == "range":
            for bs in r[1]:
                self._bytes(a, bs, b)
        elif tag == "cat":
            m = self.new()
            self.build(r[1], a, m)
            self.build(r[2], m, b)
        elif tag == "alt":
            self.build(r[1], a, b)
            self.build(r[2], a, b)
        elif tag == "star":
            m = self.new()
            self.eps[a].append(m)
            self.eps[m].append(b)
            self.build(r[1], m, m)
        else:
            raise ValueError(tag)

    def _bytes(self, a: int, bs: bytes, b: int):
        cur = a
        for j, byte in enumerate(bs):
            nxt = b if j == len(bs) - 1 else self.new()
            self.edge(cur, byte, nxt)
            cur = nxt

    def closure(self, states) -> FrozenSet[int]:
        stack = list(states)
        seen = set(states)
        while stack:
            s = stack.pop()
            for t in self.eps[s]:
                if t not in seen:
                    seen.add(t)
                    stack.append(t)
        return frozenset(seen)


@dataclasses.dataclass
class TokenDFA:
    """Partial DFA with producing transitions.

    States are 1..n (0 is the implicit dead state), 1 is the initial state.
    ``delta[s][b]`` is the next state (0 = none/dead) and ``produce[s][b]``
    whether that transition ends a token.  ``token[s]`` is the terminal key
    of an accepting state, else None.
    """

    n: int
    delta: List[List[int]]
    produce: List[List[bool]]
    token: List[Optional[tuple]]

    def accepting(self, s: int) -> bool:
        return self.token[s] is not None


def build_token_dfa(rules: Sequence[Tuple[tuple, tuple]], order: Dict[tuple, int]) -> TokenDFA:
    """``rules``: (terminal key, regex) pairs.  ``order``: terminal -> priority."""
    nfa = _NFA()
    start = nfa.new()
    finals: Dict[int, tuple] = {}
    for term, regex in rules:
        a, b = nfa.new(), nfa.new()
        nfa.eps[start].append(a)
        nfa.build(regex, a, b)
        finals[b] = term

    def tok_of(S) -> Optional[tuple]:
        best = None
        for s in S:
            t = finals.get(s)
            if t is not None and (best is None or order[t] < order[best]):
                best = t
        return best

    # Subset construction.  Only non-empty subsets that can still reach an
    # accepting state are kept (DFA.hs:150-213 drops useless states).
    init = nfa.closure([start])
    index = {init: 1}
    sets = [None, init]
    raw: List[Dict[int, FrozenSet[int]]] = [{}, {}]
    work = [init]
    while work:
        S = work.pop()
        moves: Dict[int, set] = {}
        for s in S:
            for byte, ts in nfa.trans[s].items():
                moves.setdefault(byte, set()).update(ts)
        out = {}
        for byte, ts in moves.items():
            T = nfa.closure(ts)
            if T not in index:
                index[T] = len(sets)
                sets.append(T)
                raw.append({})
                work.append(T)
            out[byte] = T
        raw[index[S]] = out
    n = len(sets) - 1
    tokens = [None] + [tok_of(S) for S in sets[1:]]
    succ = [[0] * 256 for _ in range(n + 1)]
    for i in range(1, n + 1):
        for byte, T in raw[i].items():
            succ[i][byte] = index[T]
    # Useful states: those from which an accepting state is reachable.
    useful = set(i for i in range(1, n + 1) if tokens[i] is not None)
    changed = True
    while changed:
        changed = False
        for i in range(1, n + 1):
            if i not in useful and any(succ[i][b] in useful for b in range(256) if succ[i][b]):
                useful.add(i)
                changed = True
    for i in range(1, n + 1):
        for b in range(256):
            if succ[i][b] and succ[i][b] not in useful:
                succ[i][b] = 0
    # (the initial state is always useful: every terminal is non-empty)

    produce = [[False] * 256 for _ in range(n + 1)]
    for q in range(1, n + 1):
        if q not in useful or tokens[q] is None:
            continue
        for b in range(256):
            if succ[q][b] == 0 and succ[1][b] != 0:
                succ[q][b] = succ[1][b]
                produce[q][b] = True

    # Keep only reachable useful states, renumber (initial stays 1).
    keep = [1]
    seen = {1}
    for s in keep:
        for b in range(256):
            t = succ[s][b]
            if t and t not in seen:
                seen.add(t)
                keep.append(t)
    ren = {old: new for new, old in enumerate(keep, start=1)}
    delta = [[0] * 256] + [[ren.get(succ[o][b], 0) for b in range(256)] for o in keep]
    prod = [[False] * 256] + [list(produce[o]) for o in keep]
    tok = [None] + [tokens[o] for o in keep]
    dfa = TokenDFA(len(keep), delta, prod, tok)
    return minimise(dfa)


def minimise(dfa: TokenDFA) -> TokenDFA:
    """Moore partition refinement on (token, next block, produce flag)."""
    n = dfa.n
    block = [0] * (n + 1)
    keys = {}
    for s in range(1, n + 1):
        block[s] = keys.setdefault(("t", dfa.token[s]), len(keys) + 1)
    while True:
        keys = {}
        nb = [0] * (n + 1)
        for s in range(1, n + 1):
            sig = (block[s],
                   tuple(block[dfa.delta[s][b]] if dfa.delta[s][b] else 0 for b in range(256)),
                   tuple(dfa.produce[s]))
            nb[s] = keys.setdefault(sig, len(keys) + 1)
        if len(keys) == len(set(block[1:])):
            block = nb
            break
        block = nb
    # Renumber so the initial state's block is 1, in BFS order.
    order = [block[1]]
    seen = {block[1]}
    rep = {}
    for s in range(1, n + 1):
        rep.setdefault(block[s], s)
    for blk in order:
        s = rep[blk]
        for b in range(256):
            t = dfa.delta[s][b]
            if t and block[t] not in seen:
                seen.add(block[t])
                order.append(block[t])
    ren = {blk: i for i, blk in enumerate(order, start=1)}
    m = len(order)
    delta = [[0] * 256 for _ in range(m + 1)]
    prod = [[False] * 256 for _ in range(m + 1)]
    tok = [None] * (m + 1)
    for blk in order:
        s = rep[blk]
        i = ren[blk]
        tok[i] = dfa.token[s]
        for b in range(256):
            t = dfa.delta[s][b]
            delta[i][b] = ren[block[t]] if t else 0
            prod[i][b] = dfa.produce[s][b]
    return TokenDFA(m, delta, prod, tok)


def tokenize(dfa: TokenDFA, data: bytes) -> Optional[List[Tuple[tuple, int, int]]]:
    """Direct greedy simulation of the token DFA (restates ``tokenize``,
    src/Alpacc/Lexer/DFA.hs:404-440): list of (terminal, start, end) including
    ignore tokens, or None if the input is rejected."""
    if not data:
        return []
    out = []
    s = 1
    start = 0
    n = len(data)
    for i, b in enumerate(data):
        t = dfa.delta[s][b]
        if t == 0:
            return None
        if dfa.produce[s][b]:
            out.append((dfa.token[s], start, i))
            start = i
        s = t
    if dfa.token[s] is None:
        return None
    out.append((dfa.token[s], start, n))
    return out


# --------------------------------------------------------------------------
# Transformation semigroup (the parallel lexer)
# --------------------------------------------------------------------------

IDENTITY_E = 0
DEAD_E = 1


@dataclasses.dataclass
class Endos:
    """The closed set of byte-string endomorphisms.

    ``funcs[e]``  -- tuple of next states for states 0..n (0 = dead)
    ``prods[e]``  -- tuple of produce flags of the last transition
    ``byte_endo[b]`` -- endomorphism id of byte b
    ``compose[a][b]`` -- id of "apply a, then b"
    Ids 0 and 1 are the identity and the all-dead endomorphism.
    """

    funcs: List[Tuple[int, ...]]
    prods: List[Tuple[bool, ...]]
    byte_endo: List[int]
    compose: List[List[int]]

    @property
    def size(self) -> int:
        return len(self.funcs)


def endomorphisms(dfa: TokenDFA) -> Endos:
    n = dfa.n
    ident = (tuple(range(n + 1)), tuple([False] * (n + 1)))
    dead = (tuple([0] * (n + 1)), tuple([False] * (n + 1)))
    elems = [ident, dead]
    index = {ident: 0, dead: 1}

    def intern(e) -> int:
        if e[0] == dead[0]:  # every all-dead map behaves as the dead element
            return DEAD_E
        j = index.get(e)
        if j is None:
            j = len(elems)
            index[e] = j
            elems.append(e)
        return j

    byte_endo = []
    gens = []
    for b in range(256):
        f = tuple(dfa.delta[s][b] if s else 0 for s in range(n + 1))
        p = tuple(dfa.produce[s][b] if s else False for s in range(n + 1))
        j = intern((f, p))
        byte_endo.append(j)
        if j not in gens and j != DEAD_E:
            gens.append(j)

    def comp(a, b):
        fa, _ = a
        fb, pb = b
        return (tuple(fb[x] for x in fa), tuple(pb[x] for x in fa))

    # Closure under right multiplication by generators = the semigroup.
    i = 2
    while i < len(elems):
        e = elems[i]
        for g in gens:
            intern(comp(e, elems[g]))
        i += 1
    N = len(elems)
    table = [[0] * N for _ in range(N)]
    for a in range(N):
        for b in range(N):
            if a == IDENTITY_E:
                table[a][b] = b
            elif b == IDENTITY_E:
                table[a][b] = a
            elif a == DEAD_E or b == DEAD_E:
                table[a][b] = DEAD_E
            else:
                c = comp(elems[a], elems[b])
                j = index.get(c)
                if j is None:
                    j = intern(c)
                    assert j == DEAD_E, "semigroup not closed"
                table[a][b] = j
    return Endos([e[0] for e in elems], [e[1] for e in elems], byte_endo, table)
