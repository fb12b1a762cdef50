"""LLP(q,k) parsing table (Vagner & Melichar's algorithm 8 as alpacc builds it).

Restates src/Alpacc/Parser/LLP.hs (item collection, shortest stack prefixes,
start pairs, bracket homomorphisms) over the k-sets of
src/Alpacc/Parser/LL.hs (FIRST_k, FOLLOW_k, and LAST_q / BEFORE_q as FIRST /
FOLLOW of the reversed grammar).

Grammar symbols are plain ints: ``0 <= s < n_terms`` is a terminal, anything
larger a nonterminal.  A production is (lhs, rhs-tuple).  Terminal strings
are tuples of terminal ids.
"""

from __future__ import annotations

from typing import Dict, FrozenSet, List, Optional, Sequence, Set, Tuple

Sym = int
Prod = Tuple[int, Tuple[int, ...]]


class NotLLP(ValueError):
    pass


def _tprod(k: int, A: FrozenSet[tuple], B: FrozenSet[tuple]) -> FrozenSet[tuple]:
    """Truncated product: {take k (a ++ b)} (LL.hs:173-179)."""
    out = set()
    for a in A:
        if len(a) >= k:
            out.add(a[:k])
            continue
        for b in B:
            out.add((a + b)[:k])
    return frozenset(out)


class KSets:
    """FIRST_k over one grammar, memoised on symbol strings (LL.hs:216-293)."""

    def __init__(self, k: int, n_terms: int, prods: Sequence[Prod], nonterminals: Sequence[int]):
        self.k = k
        self.n_terms = n_terms
        self.prods = list(prods)
        first = {nt: frozenset() for nt in nonterminals}
        while True:
            new = dict(first)
            for lhs, rhs in self.prods:
                new[lhs] = new[lhs] | self._first(rhs, first)
            if new == first:
                break
            first = new
        self.first_map = first
        self.memo: Dict[tuple, FrozenSet[tuple]] = {}

    def _first(self, w, fmap) -> FrozenSet[tuple]:
        acc = frozenset([()])
        for s in w:
            if s < self.n_terms:
                f = frozenset([(s,)[: self.k]])
            else:
                f = fmap[s]
            acc = _tprod(self.k, acc, f)
            if not acc:
                break
        return acc

    def first(self, w: tuple) -> FrozenSet[tuple]:
        r = self.memo.get(w)
        if r is None:
            r = self._first(w, self.first_map)
            self.memo[w] = r
        return r


def follow_sets(k: int, n_terms: int, prods: Sequence[Prod], nonterminals: Sequence[int],
                start: int) -> Dict[int, FrozenSet[tuple]]:
    """FOLLOW_k via the End-padded extended grammar (LL.hs:326-420).

    End markers are stripped from the result, so strings near the end of the
    input are shorter than k."""
    END = n_terms  # a fresh terminal id for the padding symbol
    ext_terms = n_terms + 1
    shift = 1  # nonterminal ids move up by one to make room for END

    def sh(s):
        return s + shift if s >= n_terms else s

    eprods = [(lhs + shift, tuple(sh(s) for s in rhs)) for lhs, rhs in prods]
    ents = [nt + shift for nt in nonterminals]
    ks = KSets(k, ext_terms, eprods, ents)
    follow = {nt: frozenset() for nt in ents}
    follow[start + shift] = frozenset([tuple([END] * k)])
    right = []
    for lhs, rhs in eprods:
        for i, s in enumerate(rhs):
            if s >= ext_terms:
                right.append((lhs, s, rhs[i + 1:]))
    while True:
        new = dict(follow)
        for lhs, a, rest in right:
            new[a] = new[a] | _tprod(k, ks.first(rest), follow[lhs])
        if new == follow:
            break
        follow = new
    out = {}
    for nt in nonterminals:
        out[nt] = frozenset(tuple(t for t in w if t != END) for w in follow[nt + shift])
    return out


class LLPGrammar:
    """An augmented grammar ready for LLP table construction."""

    def __init__(self, n_terms: int, prods: Sequence[Prod], nonterminals: Sequence[int],
                 start: int, q: int, k: int):
        if q < 0:
            raise NotLLP("Error: Lookback must be non-negative.")
        if k < 1:
            raise NotLLP("Error: Lookahead must be positive.")
        self.n_terms = n_terms
        self.prods = list(prods)
        self.nts = list(nonterminals)
        self.start = start
        self.q, self.k = q, k
        self.by_lhs: Dict[int, List[int]] = {}
        for i, (lhs, _) in enumerate(self.prods):
            self.by_lhs.setdefault(lhs, []).append(i)
        self.firstk = KSets(k, n_terms, self.prods, self.nts)
        self.follow = follow_sets(k, n_terms, self.prods, self.nts, start)
        rprods = [(lhs, tuple(reversed(rhs))) for lhs, rhs in self.prods]
        self.lastq_rev = KSets(q, n_terms, rprods, self.nts)
        bef = follow_sets(q, n_terms, rprods, self.nts, start)
        self.before = {nt: frozenset(tuple(reversed(w)) for w in ws) for nt, ws in bef.items()}
        self._last_memo: Dict[tuple, FrozenSet[tuple]] = {}
        self.ll_table = self._ll_table()

    def is_term(self, s: int) -> bool:
        return s < self.n_terms

    def first(self, w: tuple) -> FrozenSet[tuple]:
        return self.firstk.first(w)

    def last(self, w: tuple) -> FrozenSet[tuple]:
        r = self._last_memo.get(w)
        if r is None:
            r = frozenset(tuple(reversed(x)) for x in self.lastq_rev.first(tuple(reversed(w))))
            self._last_memo[w] = r
        return r

    def _ll_table(self) -> Dict[Tuple[int, tuple], int]:
        """LL(k) table (LL.hs:538-558); conflicts are an error."""
        table: Dict[Tuple[int, tuple], Set[int]] = {}
        for i, (lhs, rhs) in enumerate(self.prods):
            for y in _tprod(self.k, self.first(rhs), self.follow[lhs]):
                table.setdefault((lhs, y), set()).add(i)
        bad = {key: v for key, v in table.items() if len(v) > 1}
        if bad:
            raise NotLLP(f"LL({self.k}) Table Conflicts: {sorted(bad.items())[:4]}")
        return {key: min(v) for key, v in table.items()}

    # -- LLP.hs:476-492 -------------------------------------------------------
    def parse_config(self, y: tuple, alpha: tuple):
        """Run the LL(k) parser on input ``y`` with stack ``alpha`` until the
        first terminal is matched; returns (alpha, omega, pi) or None."""
        inp = y
        stack = list(alpha)
        parsed = []
        while True:
            if stack and self.is_term(stack[0]):
                if inp and inp[0] == stack[0]:
                    return (alpha, tuple(stack[1:]), tuple(parsed))
                return None
            if stack:
                top = stack[0]
                idx = self.ll_table.get((top, inp[: self.k]))
                if idx is None:
                    return None
                stack = list(self.prods[idx][1]) + stack[1:]
                parsed.append(idx)
                continue
            if not inp:
                return (alpha, (), tuple(parsed))
            return None

    def shortest_prefix(self, v: tuple, x_delta: tuple):
        """LLP.hs:96-112: the shortest stack prefix on which the LL(k) parse
        of ``v`` matches its first terminal."""
        for n in range(1, len(x_delta) + 1):
            r = self.parse_config(v, x_delta[:n])
            if r is not None:
                return r
        return None


# Item = (lhs, alpha_x, beta, suffix, prefix, config) with config=(alpha, omega, pi)


def _new_items(g: LLPGrammar, item) -> Set[tuple]:
    """LLP.hs:237-278 (step 3(a) of algorithm 8)."""
    Y, alpha_x, beta, _ui, vi, delta = item
    if not alpha_x:
        return set()
    alpha = alpha_x[:-1]
    uj = set()
    for b in g.before[Y]:
        uj |= g.last(tuple(b) + alpha)
    x = alpha_x[-1]
    vj = g.first((x,) + tuple(vi))
    x_delta = (x,) + delta[0]
    x_beta = (x,) + beta
    out = set()
    for u in uj:
        for v in vj:
            sp = g.shortest_prefix(v, x_delta)
            if sp is None:
                raise NotLLP("no stack prefix admits the lookahead")
            out.add((Y, alpha, x_beta, u, v, sp))
    return out


def _extra_items(g: LLPGrammar, item) -> Set[tuple]:
    """LLP.hs:284-318 (step 3(b))."""
    _, alpha_y, _, _, v, gamma = item
    if not alpha_y or g.is_term(alpha_y[-1]):
        return set()
    y = alpha_y[-1]
    out = set()
    for pi in g.by_lhs.get(y, []):
        delta = g.prods[pi][1]
        us = set()
        for b in g.before[y]:
            us |= g.last(tuple(b) + delta)
        for u in us:
            out.add((y, delta, (), u, v, gamma))
    return out


def _close_extra(g: LLPGrammar, items: Set[tuple]) -> FrozenSet[tuple]:
    items = set(items)
    frontier = list(items)
    while frontier:
        nxt = []
        for it in frontier:
            for e in _extra_items(g, it):
                if e not in items:
                    items.add(e)
                    nxt.append(e)
        frontier = nxt
    return frozenset(items)


def llp_table(g: LLPGrammar) -> Dict[Tuple[tuple, tuple], Tuple[tuple, tuple, tuple]]:
    """The LLP collection (LLP.hs:377-447) and start pairs (LLP.hs:495-516).

    Returns {(x, y): (alpha, omega, pi)} with collection entries taking
    precedence over start entries (LLP.hs:556)."""
    start_ps = g.by_lhs[g.start]
    if len(start_ps) != 1:
        raise NotLLP("start symbol must have exactly one production")
    sp = g.prods[start_ps[0]][1]
    d0 = frozenset((g.start, sp, (), l, (), ((), (), ())) for l in g.last(sp))
    table: Dict[Tuple[tuple, tuple], tuple] = {}
    seen: Set[FrozenSet[tuple]] = set()
    queue = [d0]
    qi = 0
    while qi < len(queue):
        top = queue[qi]
        qi += 1
        if top in seen:
            continue
        seen.add(top)
        for it in top:
            alpha_z = it[1]
            if alpha_z and g.is_term(alpha_z[-1]):
                key = (it[3], it[4])
                val = it[5]
                old = table.get(key)
                if old is not None and old != val:
                    raise NotLLP(f"not LLP({g.q},{g.k}): pair {key} has two stack configurations")
                table[key] = val
        groups: Dict[int, List[tuple]] = {}
        for it in top:
            if it[1]:
                groups.setdefault(it[1][-1], []).append(it)
        for grp in groups.values():
            new = set()
            for it in grp:
                new |= _new_items(g, it)
            nxt = _close_extra(g, new)
            if nxt and nxt not in seen:
                queue.append(nxt)
    starts = {((), y): ((g.start,), sp[1:], ()) for y in g.first(sp)}
    return table, starts


def bracket_table(table, starts):
    """Bracket-homomorphism form of the table (LLP.hs:545-565).

    Collection pairs pop their consumed stack prefix alpha (right brackets,
    top first) and push omega bottom-first (left brackets); start pairs only
    push.  Collection entries win over start entries on equal keys."""
    out = {}
    for key, (_alpha, omega, pi) in starts.items():
        out[key] = ([("L", s) for s in reversed(omega)], list(pi))
    for key, (alpha, omega, pi) in table.items():
        out[key] = ([("R", s) for s in alpha] + [("L", s) for s in reversed(omega)], list(pi))
    return out
