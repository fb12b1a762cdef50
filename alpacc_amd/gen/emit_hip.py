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
  0xffff = no pair) over the padded q+k window, or, when that index would
  exceed MAX_WINDOW_ENTRIES, an open-addressing table of the windows that
  have a pair (``WINDOW_KEYS[slot]`` the window code, all ones when empty,
  ``WINDOW[slot]`` its pair id; slot = the code times WINDOW_HASH_MUL, top
  bits; linear probing), pair spans into the bracket and production strings
  (brackets as u16: symbol | left << 15), arities and the
  terminal-production flags.
"""

from __future__ import annotations

from typing import List

from .tables import Grammar

MAX_WINDOW_ENTRIES = 1 << 24
WINDOW_HASH_MUL = 0x9E3779B97F4A7C15  # Fibonacci hashing of window codes
EMPTY_KEY = (1 << 64) - 1


def window_slot(code: int, bits: int) -> int:
    """First slot of a window code in a hashed window table of 2^bits slots."""
    return ((code * WINDOW_HASH_MUL) & EMPTY_KEY) >> (64 - bits)


class EmitError(ValueError):
    pass


def _arr(ctype: str, name: str, vals: List[int], per_line: int = 24) -> str:
    sfx = "ull" if ctype == "uint64_t" else ""
    vals = [str(int(v)) + sfx for v in vals] or ["0"]
    body = ",\n".join("  " + ", ".join(vals[i:i + per_line]) for i in range(0, len(vals), per_line))
    return f"static const {ctype} {name}[{len(vals)}] = {{\n{body}\n}};\n"


def device_tables(g: Grammar, hashed=None) -> dict:
    """Plain-int device tables (shared by the emitter and the tests).
    hashed: the window table as an open-addressing table (True), a dense
    index (False), or (None) hashed only when the dense index would exceed
    MAX_WINDOW_ENTRIES."""
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

    d = dict(
        nstate=nstate, nendo=ne, nclass=nc, byte_class=classes, step=step, compose=compose,
        endo_init=endo_init, delta_lex=delta(lo), delta_both=delta(lb), token_lex=tokens(lo), token_both=tokens(lb),
        ignore_lex=-1 if lo.ignore_id is None else lo.ignore_id,
        ignore_both=-1 if lb.ignore_id is None else lb.ignore_id,
        dead_lex=lo.dead_id, dead_both=lb.dead_id,
    )
    d.update(_parser_tables(g, hashed) if g.has_parser else _no_parser(g))
    return d


def _no_parser(g: Grammar) -> dict:
    """Parser fields of a lexer-only grammar: no windows, no pairs; record 0
    (the empty record) and one production slot keep every array non-empty."""
    return dict(
        has_parser=0, empty_t=g.lex_both.dead_id, start_t=g.lex_both.dead_id + 1, end_t=g.lex_both.dead_id + 2,
        window=[], base=g.lex_both.dead_id + 3, npairs=0, spans=[], stacks=[], productions=[], n_productions=0,
        production_bytes=1, arity=[], prod_is_term=[], profile=[], max_slot_items=1, pair_rec=[0] * 64,
        prod_info=[], max_brackets=0, max_prods=0, max_bpop=0, max_bpush=0, max_spop=0, max_spush=0, max_nt=0,
        window_hash=0, window_keys=[], pair_prod16=[],
    )


def _parser_tables(g: Grammar, hashed=None) -> dict:
    p = g.parser
    if p.production_bytes > 2:
        raise EmitError(f"{g.name}: {p.n_productions} productions exceed the u16 production field")
    base = p.end_t + 1
    w = p.q + p.k
    nwin = base ** w
    if hashed is None:
        hashed = nwin > MAX_WINDOW_ENTRIES
    if nwin >= EMPTY_KEY:
        raise EmitError(f"{g.name}: window codes of {w} digits base {base} exceed 64 bits")
    codes = []
    for key in p.keys:
        code = 0
        for t in key:
            code = code * base + t
        codes.append(code)
    window = [] if hashed else [0xFFFF] * nwin
    for i, code in enumerate(codes):
        if not hashed:
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
    bad_pairs = set()
    # two-byte production ids (more than 255 productions): the records keep
    # the low bytes, and a side table the whole ids, 8 a pair
    prod16 = []
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
            bad_pairs.add(i)
            if not hashed:
                window[codes[i]] = 0xFFFF
        eff["bpop"] = max(eff["bpop"], len(ext))
        eff["bpush"] = max(eff["bpush"], len(st))
        eff["spop"] = max(eff["spop"], spop)
        eff["spush"] = max(eff["spush"], len(slots))
        eff["nt"] = max(eff["nt"], terms)
        br = ext + [0] * (4 - len(ext)) + st + [0] * (4 - len(st))
        rec += [len(ext), len(st), len(prods), terms, spop, len(slots), tmask, 0]
        rec += [b for v in br for b in (v & 0xFF, v >> 8)]
        rec += [x & 0xFF for x in prods] + [0] * (8 - len(prods))
        prod16 += prods + [0] * (8 - len(prods))
        rec += parent + [0] * (8 - len(parent))
        rec += slots + [0] * (24 - len(slots))
    rec += [0] * 64  # record npairs: no effect (positions past an input's end)
    prod16 += [0] * 8
    if p.production_bytes == 1:
        prod16 = []
    window_hash, window_keys = 0, []
    if hashed:
        # the windows with a pair id, at most half the slots filled
        live = [(c, i) for i, c in enumerate(codes) if i not in bad_pairs]
        bits = max(4, (2 * len(live) - 1).bit_length())
        if bits > 30:
            raise EmitError(f"{g.name}: {len(live)} windows exceed the hashed window table")
        mask = (1 << bits) - 1
        window_keys = [EMPTY_KEY] * (1 << bits)
        window = [0xFFFF] * (1 << bits)
        for c, i in live:
            slot = window_slot(c, bits)
            while window_keys[slot] != EMPTY_KEY:
                slot = (slot + 1) & mask
            window_keys[slot] = c
            window[slot] = i
        window_hash = bits
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
        has_parser=1, empty_t=p.empty_t, start_t=p.start_t, end_t=p.end_t,
        window=window, base=base, npairs=len(p.keys), spans=spans, stacks=stacks16, productions=p.productions,
        n_productions=p.n_productions, production_bytes=p.production_bytes,
        arity=p.arity, prod_is_term=[1 if t is not None else 0 for t in p.prod_terminal],
        profile=profile, max_slot_items=max_slot_items, pair_rec=rec, prod_info=prod_info,
        max_brackets=p.max_brackets, max_prods=p.max_prods,
        max_bpop=eff["bpop"], max_bpush=eff["bpush"], max_spop=eff["spop"], max_spush=eff["spush"],
        max_nt=eff["nt"], window_hash=window_hash, window_keys=window_keys, pair_prod16=prod16,
    )



# Fields of GrammarDesc (csrc/grammar.hpp) in declaration order: a scalar
# ("int"), the name ("str"), or an array of u8 / u16 / i32 entries.  The
# C++ source and the binary tables file are both written from this list.
def _fields(g: Grammar, d: dict) -> list:
    def one(v):  # an array is never empty (the C++ arrays hold at least one entry)
        return list(v) or [0]
    return [
        ("name", "str", g.name), ("q", "int", g.q), ("k", "int", g.k),
        ("nstate", "int", d["nstate"]), ("nendo", "int", d["nendo"]), ("nclass", "int", d["nclass"]),
        ("byte_class", "u8", d["byte_class"]), ("step", "u16", d["step"]), ("compose", "u16", d["compose"]),
        ("endo_init", "u8", d["endo_init"]), ("delta_lex", "u16", d["delta_lex"]),
        ("delta_both", "u16", d["delta_both"]), ("token_lex", "u8", d["token_lex"]),
        ("token_both", "u8", d["token_both"]),
        ("ignore_lex", "int", d["ignore_lex"]), ("ignore_both", "int", d["ignore_both"]),
        ("dead_lex", "int", d["dead_lex"]), ("dead_both", "int", d["dead_both"]),
        ("terminal_bytes_lex", "int", g.terminal_bytes_lex), ("terminal_bytes_both", "int", g.terminal_bytes_both),
        ("index_bytes", "int", g.index_bytes), ("length_bytes", "int", g.length_bytes),
        ("empty_t", "int", d["empty_t"]), ("start_t", "int", d["start_t"]), ("end_t", "int", d["end_t"]),
        ("window_base", "int", d["base"]),
        ("nwindow", "int", len(d["window"])), ("window", "u16", one(d["window"])),
        ("npairs", "int", d["npairs"]), ("spans", "i32", one(d["spans"])), ("profile", "u8", one(d["profile"])),
        ("nstacks", "int", max(1, len(d["stacks"]))), ("stacks", "u16", one(d["stacks"])),
        ("nprods_flat", "int", max(1, len(d["productions"]))), ("productions", "u16", one(d["productions"])),
        ("nproductions", "int", d["n_productions"]), ("production_bytes", "int", d["production_bytes"]),
        ("arity", "u8", one(d["arity"])), ("prod_is_terminal", "u8", one(d["prod_is_term"])),
        ("max_brackets", "int", max(1, d["max_brackets"])), ("max_prods", "int", max(1, d["max_prods"])),
        ("max_slot_items", "int", d["max_slot_items"]),
        ("pair_rec", "u8", d["pair_rec"]), ("prod_info", "u8", one(d["prod_info"])),
        ("max_bpop", "int", max(1, d["max_bpop"])), ("max_bpush", "int", max(1, d["max_bpush"])),
        ("max_spop", "int", max(1, d["max_spop"])), ("max_spush", "int", max(1, d["max_spush"])),
        ("max_nt", "int", max(1, d["max_nt"])),
        ("length_param", "int", g.length_bits or 0),
        ("has_parser", "int", d["has_parser"]),
        ("terminal_names_lex", "u8", _names(g.lex_only.names())),
        ("terminal_names_both", "u8", _names(g.lex_both.names())),
        ("production_names", "u8", _names(g.parser.prod_names if g.has_parser else [])),
        ("window_hash", "int", d["window_hash"]), ("window_keys", "u64", one(d["window_keys"])),
        ("pair_prod16", "u16", one(d["pair_prod16"])),
    ]


def _names(names) -> list:
    """Names as one NUL-separated byte array (at least one NUL)."""
    return list(b"".join(n.encode() + b"\0" for n in names)) or [0]


_CTYPE = {"u8": "uint8_t", "u16": "uint16_t", "i32": "int32_t", "u64": "uint64_t"}


def emit(g: Grammar) -> str:
    """C++ source of the grammar's GrammarDesc (compiled into the library)."""
    d = device_tables(g)
    fields = _fields(g, d)
    ns = f"grammar_{g.name}"
    src = [f"// generated by alpacc_amd.gen.emit_hip for grammar '{g.name}'; do not edit",
           '#include "../grammar.hpp"', "namespace alpacc {", f"namespace {ns} {{"]
    init = []
    for name, kind, v in fields:
        if kind == "str":
            init.append(f'"{v}"')
        elif kind == "int":
            init.append(str(int(v)))
        else:
            arr = _arr(_CTYPE[kind], name.upper(), v)
            src.append(("__attribute__((aligned(64)))\n" if name == "pair_rec" else "") + arr)
            init.append(f"{ns}::{name.upper()}")
    src += ["}  // namespace", "", f"extern const GrammarDesc GRAMMAR_{g.name.upper()} = {{"]
    src += ["  " + ", ".join(init[i:i + 6]) + "," for i in range(0, len(init), 6)]
    src += ["};", "}  // namespace alpacc", ""]
    return "\n".join(src)


TABLES_MAGIC = b"ALPHIPT1"


def emit_tables(g: Grammar, hashed=None, name=None) -> bytes:
    """The same descriptor as a binary tables file that the library loads at
    run time (alpacc_hip_load_grammar): the magic, a u32 field count, then
    every field in order -- a string as u32 length + bytes, a scalar as
    tag 'I' + i64, an array as tag 'A' + u32 entry size + u64 count + the
    entries; each field padded to 8 bytes.  hashed: as device_tables; name:
    the grammar's name in the file (default its own)."""
    import struct
    d = device_tables(g, hashed)
    fields = _fields(g, d)
    if name is not None:
        fields[0] = ("name", "str", name)
    out = bytearray(TABLES_MAGIC)
    out += struct.pack("<I", len(fields))

    def pad():
        out.extend(b"\0" * (-len(out) % 8))

    pad()
    for name, kind, v in fields:
        if kind == "str":
            b = v.encode()
            out += struct.pack("<I", len(b)) + b
        elif kind == "int":
            out += b"I" + b"\0" * 7 + struct.pack("<q", int(v))
        else:
            fmt, size = {"u8": ("B", 1), "u16": ("H", 2), "i32": ("i", 4), "u64": ("Q", 8)}[kind]
            out += b"A" + b"\0" * 3 + struct.pack("<IQ", size, len(v)) + struct.pack(f"<{len(v)}{fmt}", *v)
        pad()
    return bytes(out)
