"""Grammar diagnostics of the lexer's endomorphisms: the counterparts of
`alpacc dev composition-histogram | image-sizes | raw-image-sizes |
dump-endo` (app/Main.hs:645-815, src/Alpacc/Analysis/CompositionHistogram.hs).

The reports are computed from this generator's token DFA and endomorphism
closure.  Endomorphism ids are this generator's own (the reference's
numbering never reaches any output either), so "encoded" values differ from
the reference's; the counts and sizes do not depend on the numbering.
"""

from __future__ import annotations

from collections import Counter
from typing import List, Tuple

from .tables import Grammar


def _report_lines(title: str, per_char, closure) -> List[str]:
    max_img = max([i for _, i, _ in per_char] + [0])
    max_pairs = max([p for _, _, p in per_char] + [0])
    lines = [title,
             f"Per-char endomorphisms: {len(per_char)}",
             f"  max distinct non-dead image states (per-char) : {max_img}",
             f"  max non-dead (in,out) pair count   (per-char) : {max_pairs}",
             f"Monoid closure size                            : {len(closure)}",
             f"  max distinct non-dead image states (closure) : {max([_image(f) for f in closure] + [0])}",
             f"  max non-dead (in,out) pair count   (closure) : {max([_pairs(f) for f in closure] + [0])}",
             "", "Top per-char image sizes (first 20 by pair count):"]
    for c, i, p in sorted(per_char, key=lambda x: -x[2])[:20]:
        lines.append(f"  {c} (0x{c:02x}): image={i} pairs={p}")
    return lines


# An endomorphism is (next state of every state, produce flag of every
# state's transition), as the reference's Endomorphism carries both; state 0
# is dead.
def _image(e) -> int:
    """Distinct non-dead states in the image."""
    return len({j for j in e[0] if j != 0})


def _pairs(e) -> int:
    return sum(1 for s, j in enumerate(e[0]) if s != 0 and j != 0)


def _closure(singles) -> set:
    """Monoid closure by breadth-first composition (maxMonoidImageSize)."""
    seen = set(singles)
    queue = list(seen)
    while queue:
        e = queue.pop(0)
        for s in singles:
            c = (tuple(s[0][x] for x in e[0]), tuple(s[1][x] for x in e[0]))  # e then s
            if c not in seen:
                seen.add(c)
                queue.append(c)
    return seen


def _byte_maps(g: Grammar, strip_produce: bool):
    """The endomorphism of every byte with a transition out of some state
    (raw: producing transitions removed)."""
    dfa = g.lex_both.dfa
    n = dfa.n + 1
    out = {}
    for b in range(256):
        f, p = [0] * n, [False] * n
        for s in range(1, n):
            prod = bool(dfa.produce[s][b])
            f[s] = 0 if (strip_produce and prod) else dfa.delta[s][b]
            p[s] = prod and not strip_produce
        if any(f):
            out[b] = (tuple(f), tuple(p))
    return out


def image_sizes(g: Grammar, raw: bool = False) -> str:
    """image-sizes (Main.hs:658-711) or, raw, raw-image-sizes (:712-769):
    per-byte DFA maps (raw: producing transitions removed) and their
    composition closure."""
    maps = _byte_maps(g, raw)
    per_char = [(b, _image(f), _pairs(f)) for b, f in sorted(maps.items())]
    title = ("Raw-DFA endomorphisms (produce-extension pairs stripped):" if raw else "Extended-DFA endomorphisms:")
    return "\n".join(_report_lines(title, per_char, _closure(list(set(maps.values()))))) + "\n"


def dump_endo(g: Grammar, ch: int) -> str:
    """dump-endo (Main.hs:770-815): the map of one byte, raw and added pairs."""
    dfa = g.lex_both.dfa
    e = _byte_maps(g, False).get(ch)
    if e is None:
        return f"Char {ch} (0x{ch:02x}) not in endo table\n"
    f = e[0]
    nd = [(s, j) for s, j in enumerate(f) if s and j]
    raw = [(s, j) for s, j in nd if not dfa.produce[s][ch]]
    added = [(s, j) for s, j in nd if dfa.produce[s][ch]]
    lines = [f"Char {ch} (0x{ch:02x}):",
             f"  total states                    : {len(f)}",
             f"  non-dead (src -> dst) pair count (extended DFA) : {len(nd)}",
             f"  distinct sources (extended)     : {len({s for s, _ in nd})}",
             f"  distinct destinations (extended): {len({j for _, j in nd})}",
             f"  raw DFA pair count              : {len(raw)}",
             f"  raw DFA distinct sources        : {len({s for s, _ in raw})}",
             f"  raw DFA distinct destinations   : {len({j for _, j in raw})}",
             f"  produce-extension pair count    : {len(added)}",
             "  RAW pairs (src -> dst):"]
    lines += [f"    {s} -> {j}" for s, j in raw]
    lines.append("  ADDED-by-produce pairs (src -> dst):")
    lines += [f"    {s} -> {j}" for s, j in added]
    return "\n".join(lines) + "\n"


def composition_histogram(g: Grammar) -> str:
    """composition-histogram (CompositionHistogram.hs): value frequencies of
    the N x N composition table, top 32."""
    lt = g.lex_both
    n = lt.endos.size
    freq = Counter(lt.compose)
    packed_ident = lt.identity
    # the dead endomorphism maps every state to the dead state 0
    dead_e = next((e for e in range(n) if all(x == 0 for x in lt.endos.funcs[e][1:])), None)
    dead = None if dead_e is None else lt.compose[dead_e * n + dead_e]
    lines = ["Composition-table value histogram",
             f"  N (endomorphisms)   : {n}",
             f"  Total cells (N*N)   : {n * n}",
             f"  dead endo (encoded) : {dead}",
             f"  identity encoded    : {packed_ident}",
             "",
             "Top 32 output values by cell count:",
             "  rank  value          cells      %      cum%   note"]
    cum = 0
    for rank, (v, cells) in enumerate(sorted(freq.items(), key=lambda x: -x[1])[:32], 1):
        cum += cells
        note = "dead" if v == dead else ("identity" if v == packed_ident else "")
        lines.append(f"  {rank:<6}{v:<15}{cells:<11}{100 * cells // (n * n)}%".ljust(39)
                     + f"{100 * cum // (n * n)}%".ljust(7) + note)
    return "\n".join(lines) + "\n"
