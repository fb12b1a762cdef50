"""Emit the device tables of one grammar as a C++ source for the HIP library.

The GPU path does not use the oracle's packed-endomorphism format.  It needs:

* ``BYTE_CLASS[256]`` -- bytes with equal endomorphisms share a class;
* ``STEP[e * NCLASS + c]`` -- endomorphism e followed by class c (u16), the
  per-byte fold of the first lexer pass (staged in LDS);
* ``COMPOSE[a * NENDO + b]`` -- "a then b" for the scans (u16, L2-resident);
* ``ENDO_INIT[e]`` -- the DFA state e maps the initial state to;
* ``DELTA_*[s * 256 + b]`` -- the token DFA itself, u16 packed as
  ``next (8 bits) | produce << 8 | kept << 9`` (kept: the closed token is
  not ignored) and ``TOKEN_*[s]`` the terminal state s accepts, for the two
  terminal numberings (lexer-only, combined), staged in LDS for the lexer's
  last pass;
* the LLP pair table as a dense window index (``WINDOW[key] -> pair id``,
  0xffff = no pair) over the padded q+k window, pair spans into the bracket
  and production strings (brackets as u16: symbol | left << 15), arities and
  the terminal-production flags.
"""

from __future__ import annotations

from typing import List

from .tables import Grammar

MAX_WINDOW_ENTRIES = 1 << 24


class EmitError(ValueError):
    pass


def _arr(ctype: str, name: str, vals: List[int], per_line: int = 24) -> str:
    vals = [str(int(v)) for v in vals] or ["0"]
    body = ",\n".join("  " + ", ".join(vals[i:i + per_line]) for i in range(0, len(vals), per_line))
    return f"static const {ctype} {name}[{len(vals)}] = {{\n{body}\n}};\n"


def device_tables(g: Grammar) -> dict:
    """Plain-int device tables (shared by the emitter and the tests)."""
    lb, lo = g.lex_both, g.lex_only
    endos = lb.endos
    dfa = lb.dfa
    nstate = dfa.n + 1
    if nstate > 256:
        raise EmitError(f"{g.name}: {nstate} DFA states exceed the u8 state field")
    classes: List[int] = []
    class_of_endo = {}
    for b in range(256):
        e = endos.byte_endo[b]
        if e not in class_of_endo:
            class_of_endo[e] = len(class_of_endo)
        classes.append(class_of_endo[e])
    class_endo = sorted(class_of_endo, key=lambda e: class_of_endo[e])
    ne, nc = endos.size, len(class_endo)
    if ne >= 0xFFFF:
        raise EmitError(f"{g.name}: {ne} endomorphisms exceed u16")
    step = [endos.compose[e][ce] for e in range(ne) for ce in class_endo]
    compose = [endos.compose[a][b] for a in range(ne) for b in range(ne)]
    endo_init = [endos.funcs[e][1] for e in range(ne)]

    def delta(lt) -> List[int]:
        # next | produce << 8 | kept << 9: a boundary before the byte, and
        # whether the token it closes (the one state s accepts) is kept
        out = []
        ign = lt.ignore_id
        for s in range(nstate):
            tok = dfa.token[s] if s else None
            t = lt.term_id[tok] if tok is not None else lt.dead_id
            for b in range(256):
                nxt = dfa.delta[s][b] if s else 0
                p = 1 if (s and dfa.produce[s][b]) else 0
                out.append(nxt | (p << 8) | ((p if t != ign else 0) << 9))
        return out

    def tokens(lt) -> List[int]:
        # the terminal state s accepts (the token a boundary after it closes)
        out = []
        for s in range(nstate):
            tok = dfa.token[s] if s else None
            t = lt.term_id[tok] if tok is not None else lt.dead_id
            if t > 255:
                raise EmitError(f"{g.name}: terminal id {t} exceeds u8")
            out.append(t)
        return out

    p = g.parser
    if p.production_bytes != 1:
        raise EmitError(f"{g.name}: {p.n_productions} productions exceed the u8 production field")
    base = p.end_t + 1
    w = p.q + p.k
    nwin = base ** w
    if nwin > MAX_WINDOW_ENTRIES:
        raise EmitError(f"{g.name}: window index of {nwin} entries is too large")
    window = [0xFFFF] * nwin
    for i, key in enumerate(p.keys):
        code = 0
        for t in key:
            code = code * base + t
        window[code] = i
    top = 1 << (8 * p.bracket_bytes - 1)
    stacks16 = [(v & (top - 1)) | (0x8000 if v & top else 0) for v in p.stacks]
    if p.n_symbols >= 0x8000:
        raise EmitError(f"{g.name}: {p.n_symbols} grammar symbols exceed the 15-bit bracket field")
    spans = []
    for (s0, s1, p0, p1) in p.spans:
        spans += [s0, s1 - s0, p0, p1 - p0]
    # Per-pair depth profiles for the tiled parser: net bracket depth, the
    # minimum of {0, depth after each bracket}; net open-child-slot depth
    # (sum of arity - 1), the minimum of {0, depth after each slot's pop};
    # the number of terminal productions.
    profile = []
    for (s0, s1, p0, p1) in p.spans:
        d = dmin = 0
        for v in stacks16[s0:s1]:
            d += 1 if v & 0x8000 else -1
            dmin = min(dmin, d)
        e = emin = 0
        nt = 0
        for prod in p.productions[p0:p1]:
            emin = min(emin, e - 1)
            e += p.arity[prod] - 1
            nt += 1 if p.prod_terminal[prod] is not None else 0
        for v in (d, dmin, e, emin):
            if not -128 <= v <= 127:
                raise EmitError(f"{g.name}: pair depth profile {v} exceeds i8")
        profile += [d & 0xFF, dmin & 0xFF, e & 0xFF, emin & 0xFF, nt, 0, 0, 0]
    # Pair records, 64 bytes each: the window's net effect.  Inside one
    # window a right bracket that meets an earlier left bracket of the same
    # window cancels it (a mismatch there rejects every input containing
    # the window, so such windows get no pair id), and a derivation slot
    # whose pop meets a child slot opened earlier in the window gets its
    # parent inside the window.  What is left: bracket pops reaching below
    # the window (in pop order), the window's surviving pushes (bottom to
    # top), slot pops reaching below it and surviving child slots.
    #   0 bracket pops   1 bracket pushes   2 productions   3 terminal productions
    #   4 slot pops      5 slot pushes      6 terminal-production mask   7 0
    #   8..15   u16 bracket symbols of the pops, 16..23 of the pushes
    #   24..31  productions
    #   32..39  parent of production j: 0x80 | e for the window's e-th slot
    #           pop, else the window-local compact index of its parent
    #   40..63  owners of the surviving child slots (window-local compact index)
    # Window-local compact index: production index minus the terminal
    # productions before it (nodes of the compact tree).
    rec = []
    eff = dict(bpop=0, bpush=0, spop=0, spush=0, nt=0)
    for i, (s0, s1, p0, p1) in enumerate(p.spans):
        st, ext = [], []
        ok = True
        for v in stacks16[s0:s1]:
            if v & 0x8000:
                st.append(v & 0x7FFF)
            elif st:
                ok &= st.pop() == v
            else:
                ext.append(v)
        prods = list(p.productions[p0:p1])
        if len(prods) > 8:
            raise EmitError(f"{g.name}: a window has more than 8 productions")
        if len(ext) > 4 or len(st) > 4:
            raise EmitError(f"{g.name}: a window keeps more than 4 bracket pops or 4 pushes")
        compact, terms, tmask = [], 0, 0
        for j, prod in enumerate(prods):
            compact.append(j - terms)
            if p.prod_terminal[prod] is not None:
                terms += 1
                tmask |= 1 << j
        slots, parent, spop = [], [], 0
        for j, prod in enumerate(prods):
            if slots:
                parent.append(slots.pop())
            else:
                parent.append(0x80 | spop)
                spop += 1
            if p.arity[prod] and tmask >> j & 1:
                raise EmitError(f"{g.name}: terminal production with children")
            slots += [compact[j]] * p.arity[prod]
        if len(slots) > 24:
            raise EmitError(f"{g.name}: a window keeps more than 24 child slots")
        if not ok:
            for code in [c for c, v in enumerate(window) if v == i]:
                window[code] = 0xFFFF
        eff["bpop"] = max(eff["bpop"], len(ext))
        eff["bpush"] = max(eff["bpush"], len(st))
        eff["spop"] = max(eff["spop"], spop)
        eff["spush"] = max(eff["spush"], len(slots))
        eff["nt"] = max(eff["nt"], terms)
        br = ext + [0] * (4 - len(ext)) + st + [0] * (4 - len(st))
        rec += [len(ext), len(st), len(prods), terms, spop, len(slots), tmask, 0]
        rec += [b for v in br for b in (v & 0xFF, v >> 8)]
        rec += prods + [0] * (8 - len(prods))
        rec += parent + [0] * (8 - len(parent))
        rec += slots + [0] * (24 - len(slots))
    rec += [0] * 64  # record npairs: no effect (positions past an input's end)
    prod_info = []
    for prod in range(p.n_productions):
        if p.arity[prod] > 127:
            raise EmitError(f"{g.name}: arity above 127")
        prod_info.append(p.arity[prod] | (0x80 if p.prod_terminal[prod] is not None else 0))
    # LDS stack room one position can need: its brackets, and its slots plus
    # the child slots they open
    max_slot_items = max([(p1 - p0) + sum(p.arity[x] for x in p.productions[p0:p1])
                          for (_, _, p0, p1) in p.spans] + [1])
    return dict(
        nstate=nstate, nendo=ne, nclass=nc, byte_class=classes, step=step, compose=compose,
        endo_init=endo_init, delta_lex=delta(lo), delta_both=delta(lb), token_lex=tokens(lo), token_both=tokens(lb),
        ignore_lex=-1 if lo.ignore_id is None else lo.ignore_id,
        ignore_both=-1 if lb.ignore_id is None else lb.ignore_id,
        dead_lex=lo.dead_id, dead_both=lb.dead_id,
        window=window, base=base, spans=spans, stacks=stacks16, productions=p.productions,
        arity=p.arity, prod_is_term=[1 if t is not None else 0 for t in p.prod_terminal],
        profile=profile, max_slot_items=max_slot_items, pair_rec=rec, prod_info=prod_info,
        max_bpop=eff["bpop"], max_bpush=eff["bpush"], max_spop=eff["spop"], max_spush=eff["spush"],
        max_nt=eff["nt"],
    )


def emit(g: Grammar) -> str:
    d = device_tables(g)
    p = g.parser
    ns = f"grammar_{g.name}"
    src = [
        f"// generated by alpacc_amd.gen.emit_hip for grammar '{g.name}'; do not edit",
        '#include "../grammar.hpp"',
        "namespace alpacc {",
        f"namespace {ns} {{",
        _arr("uint8_t", "BYTE_CLASS", d["byte_class"]),
        _arr("uint16_t", "STEP", d["step"]),
        _arr("uint16_t", "COMPOSE", d["compose"]),
        _arr("uint8_t", "ENDO_INIT", d["endo_init"]),
        _arr("uint16_t", "DELTA_LEX", d["delta_lex"]),
        _arr("uint16_t", "DELTA_BOTH", d["delta_both"]),
        _arr("uint8_t", "TOKEN_LEX", d["token_lex"]),
        _arr("uint8_t", "TOKEN_BOTH", d["token_both"]),
        _arr("uint16_t", "WINDOW", d["window"]),
        _arr("int32_t", "SPANS", d["spans"]),
        _arr("uint8_t", "PROFILE", d["profile"]),
        "__attribute__((aligned(64)))\n" + _arr("uint8_t", "PAIR_REC", d["pair_rec"]),
        _arr("uint8_t", "PROD_INFO", d["prod_info"]),
        _arr("uint16_t", "STACKS", d["stacks"]),
        _arr("uint8_t", "PRODUCTIONS", d["productions"]),
        _arr("uint8_t", "ARITY", d["arity"]),
        _arr("uint8_t", "PROD_IS_TERMINAL", d["prod_is_term"]),
        "}  // namespace",
        "",
        f"extern const GrammarDesc GRAMMAR_{g.name.upper()} = {{",
        f'  "{g.name}", {g.q}, {g.k},',
        f"  {d['nstate']}, {d['nendo']}, {d['nclass']},",
        f"  {ns}::BYTE_CLASS, {ns}::STEP, {ns}::COMPOSE, {ns}::ENDO_INIT,",
        f"  {ns}::DELTA_LEX, {ns}::DELTA_BOTH, {ns}::TOKEN_LEX, {ns}::TOKEN_BOTH,",
        f"  {d['ignore_lex']}, {d['ignore_both']}, {d['dead_lex']}, {d['dead_both']},",
        f"  {g.terminal_bytes_lex}, {g.terminal_bytes_both}, {g.index_bytes}, {g.length_bytes},",
        f"  {p.empty_t}, {p.start_t}, {p.end_t}, {d['base']},",
        f"  {len(d['window'])}, {ns}::WINDOW, {len(p.keys)}, {ns}::SPANS, {ns}::PROFILE,",
        f"  {max(1, len(p.stacks))}, {ns}::STACKS,",
        f"  {max(1, len(p.productions))}, {ns}::PRODUCTIONS,",
        f"  {p.n_productions}, {p.production_bytes}, {ns}::ARITY, {ns}::PROD_IS_TERMINAL,",
        f"  {max(1, p.max_brackets)}, {max(1, p.max_prods)}, {d['max_slot_items']},",
        f"  {ns}::PAIR_REC, {ns}::PROD_INFO,",
        f"  {max(1, d['max_bpop'])}, {max(1, d['max_bpush'])}, {max(1, d['max_spop'])}, {max(1, d['max_spush'])},",
        f"  {max(1, d['max_nt'])},",
        f"  {g.length_bits or 0},",
        "};",
        "}  // namespace alpacc",
        "",
    ]
    return "\n".join(src)
