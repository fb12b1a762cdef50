"""Pins the CPU oracle (oracle/oracle.c + the restated table generator).

* known answers: README.md:81-107 (arithmetic ``12+(3*4)``, renumbered to
  alpacc's numbering) and haskell-tests/LLP.hs:495-499 (paper grammar
  ``a+[a+a]`` for every q, k in 1..3);
* the lexer against a direct greedy DFA simulation (DFA.hs:404-440) on every
  short string over a grammar-relevant alphabet and on generated text;
* the parser against an independent LL(k) predictive parser over the same
  augmented grammar (LL.hs:615-631), on every short token string and on
  random derivations: same verdict, same derivation;
* parents against the closed form "last earlier slot with e[j] <= e[i]",
  e = exclusive scan of (arity - 1), which the GPU path computes;
* the LLP tables of the paper grammar accept every derivable string up to
  length 8 and reject every non-derivable string of length 3
  (haskell-tests/LLP.hs:505-537) for every q, k in 1..3.
"""

import itertools
import os
import random

import numpy as np
import pytest

import pyoracle
from alpacc_amd import wire
from alpacc_amd.gen import lexgen, tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def lex_layout(g, mode):
    if mode == "lexer":
        return wire.Layout("lexer", g.terminal_bytes_lex, g.index_bytes, g.length_bytes)
    if mode == "parser":
        return wire.Layout("parser", g.terminal_bytes_both, g.index_bytes, 0, g.parser.production_bytes)
    return wire.Layout("both", g.terminal_bytes_both, g.index_bytes, g.length_bytes, g.parser.production_bytes)


def oracle_rec(g, mode, payload, n=None):
    raw = pyoracle.record(g.name, mode, payload, n)
    r, off = wire.decode_record(raw, 0, lex_layout(g, mode))
    assert off == len(raw)
    return r


# --------------------------------------------------------------------------- KATs

def test_kat_arithmetic_lexer(analysed):
    g = analysed["arithmetic"]
    r = oracle_rec(g, "lexer", b"12+(3*4)")
    assert r.valid
    assert list(zip(r.terminals.tolist(), r.starts.tolist(), r.lengths.tolist())) == [
        (0, 0, 2), (5, 2, 1), (2, 3, 1), (0, 4, 1), (4, 5, 1), (0, 6, 1), (3, 7, 1)]


def test_kat_arithmetic_parser(analysed):
    g = analysed["arithmetic"]
    r = oracle_rec(g, "parser", bytes([0, 5, 2, 0, 4, 0, 3]))
    assert r.valid
    assert r.productions.tolist() == [0, 4, 9, 11, 7, 1, 17, 4, 8, 14, 0, 4, 9, 11, 5, 16, 9, 11, 7, 3, 15, 7, 3]


def test_kat_arithmetic_combined(analysed):
    g = analysed["arithmetic"]
    r = oracle_rec(g, "both", b"12+(3*4)")
    assert r.valid
    assert r.terminals.tolist() == [0, 5, 2, 0, 4, 0, 3]
    assert r.starts.tolist() == [0, 2, 3, 4, 5, 6, 7]
    assert r.lengths.tolist() == [2, 1, 1, 1, 1, 1, 1]
    assert r.nodes.tolist() == [0, 4, 9, 7, 1, 4, 8, 0, 4, 9, 5, 9, 7, 3, 7, 3]
    assert r.parents.tolist() == [0, 0, 1, 1, 0, 4, 5, 6, 7, 8, 8, 10, 10, 7, 5, 4]
    assert r.token_parents.tolist() == [2, 4, 6, 9, 10, 11, 6]
    # README.md:84-106 tree (full derivation numbering) is reproduced by the
    # parent rule on the full derivation
    full = [0, 4, 9, 11, 7, 1, 17, 4, 8, 14, 0, 4, 9, 11, 5, 16, 9, 11, 7, 3, 15, 7, 3]
    assert closed_form_parents(full, g.parser.arity) == [
        0, 0, 1, 2, 1, 0, 5, 5, 7, 8, 8, 10, 11, 12, 11, 14, 14, 16, 14, 10, 8, 7, 5]


PAPER_USER_PRODS = 5  # E->T E_, E_->+T E_, E_->, T->a, T->[E]


@pytest.mark.parametrize("q,k", [(q, k) for q in (1, 2, 3) for k in (1, 2, 3)])
def test_kat_llp_paper_all_qk(q, k):
    text = open(os.path.join(ROOT, "grammars", "paper_grammar.alp")).read()
    g = tables.analyse("paper_grammar", text, q=q, k=k)
    ids = {t[1]: i for t, i in g.lex_both.term_id.items()}
    toks = [ids[c] for c in "a+[a+a]"]
    ok, deriv = ref_llp_parse(g, toks)
    assert ok
    assert [p for p in deriv if p < PAPER_USER_PRODS] == [0, 3, 1, 4, 0, 3, 1, 3, 2, 2]
    # derivables up to length 8 parse; nondeterminables of length 3 fail
    derivable = derivable_strings(g, 8)
    for s in derivable:
        assert ref_llp_parse(g, s)[0], (q, k, s)
    terms = sorted(set(ids[c] for c in "a+[]"))
    for s in itertools.product(terms, repeat=3):
        if list(s) not in derivable:
            assert not ref_llp_parse(g, list(s))[0], (q, k, s)


# ----------------------------------------------------------- reference restatements

def ref_llp_parse(g, toks):
    """parse_test over the Python tables (windows, bracket stack)."""
    p = g.parser
    keyed = dict(zip(p.keys, p.values))
    framed = [p.start_t] + list(toks) + [p.end_t]
    m = len(framed)
    stack, deriv = [], []
    top = 1 << (8 * p.bracket_bytes - 1)
    for i in range(m):
        w = tuple(p.empty_t if (i + j < p.q or i + j >= m + p.q) else framed[i + j - p.q]
                  for j in range(p.q + p.k))
        if w not in keyed:
            return False, None
        br, ps = keyed[w]
        for b in br:
            if b & top:
                stack.append(b & ~top)
            elif not stack or stack.pop() != b:
                return False, None
        deriv += ps
    return (not stack), deriv


def ll_parse(g, toks):
    """Independent LL(k) predictive parse of ``toks`` (LL.hs:615-631) over the
    augmented grammar: stack [S, end], input tokens + [end]; returns
    (accepted, productions applied in order)."""
    llg = g.llp_grammar
    p = g.parser
    start_nt = g.productions[-1][1][1]
    inp = list(toks) + [p.end_t]
    stack = [start_nt, p.end_t]
    deriv = []
    pos = 0
    while stack:
        top = stack.pop(0)
        if llg.is_term(top):
            if pos < len(inp) and inp[pos] == top:
                pos += 1
                continue
            return False, None
        idx = llg.ll_table.get((top, tuple(inp[pos:pos + llg.k])))
        if idx is None:
            return False, None
        deriv.append(idx)
        stack = list(g.productions[idx][1]) + stack
    return pos == len(inp), deriv


def derivable_strings(g, maxlen):
    """All terminal strings of length <= maxlen derivable from the user start
    symbol (breadth-first leftmost expansion, LL.hs:48-60)."""
    p = g.parser
    prods = g.productions
    start = prods[-1][1][1]
    out = set()
    seen = set()
    frontier = [(start,)]
    nterm = p.end_t + 1
    while frontier:
        nxt = []
        for form in frontier:
            # first nonterminal
            idx = next((i for i, s in enumerate(form) if s >= nterm), None)
            if sum(1 for s in form if s < nterm) > maxlen:
                continue
            if idx is None:
                out.add(form)
                continue
            for lhs, rhs in prods:
                if lhs == form[idx]:
                    nf = form[:idx] + tuple(rhs) + form[idx + 1:]
                    if len([s for s in nf if s < nterm]) <= maxlen and len(nf) <= 3 * maxlen + 4 and nf not in seen:
                        seen.add(nf)
                        nxt.append(nf)
        frontier = nxt
    return [list(s) for s in out]


def closed_form_parents(deriv, arity):
    e = [0]
    for p in deriv[:-1]:
        e.append(e[-1] + arity[p] - 1)
    par = [0]
    for i in range(1, len(deriv)):
        j = i - 1
        while e[j] > e[i]:
            j -= 1
        par.append(j)
    return par


# ------------------------------------------------------------- lexer cross-checks

def direct_lex(g, mode, data):
    lt = g.lex_only if mode == "lexer" else g.lex_both
    toks = lexgen.tokenize(lt.dfa, data)
    if toks is None:
        return None
    maxlen = (1 << (8 * g.length_bytes)) - 1
    out = []
    for term, s, e in toks:
        tid = lt.term_id[term]
        if tid == lt.ignore_id:
            continue
        if e - s > maxlen:
            return None
        out.append((tid, s, e - s))
    return out


ALPHABETS = {
    "arithmetic": b"1-a(+ )*",
    "json": b'"a1 [,]{:}-.tn',
    "sexp": b"a1 ()\n-",
    "paper_grammar": b"1+[] ",
}


@pytest.mark.parametrize("name", sorted(ALPHABETS))
def test_lexer_exhaustive_short(analysed, name):
    g = analysed[name]
    alpha = ALPHABETS[name]
    maxn = 5 if len(alpha) > 9 else 6
    for n in range(0, maxn + 1):
        for tup in itertools.product(alpha, repeat=n):
            data = bytes(tup)
            exp = direct_lex(g, "lexer", data)
            r = oracle_rec(g, "lexer", data)
            assert r.valid == (exp is not None), data
            if exp is not None:
                got = list(zip(r.terminals.tolist(), r.starts.tolist(), r.lengths.tolist()))
                assert got == exp, data


@pytest.mark.parametrize("name", ["arithmetic", "json", "sexp"])
def test_lexer_generated_text(analysed, name):
    from alpacc_amd import datagen
    try:
        datagen._load()
    except ImportError:
        pytest.skip("data generator not built")
    g = analysed[name]
    for seed in range(4):
        for conf in (True, False):
            data = datagen.generate(name, 20000, seed, conforming=conf).tobytes()
            exp = direct_lex(g, "lexer", data)
            r = oracle_rec(g, "lexer", data)
            assert r.valid == (exp is not None)
            if exp is not None:
                got = list(zip(r.terminals.tolist(), r.starts.tolist(), r.lengths.tolist()))
                assert got == exp


def test_length_overflow_rule(analysed):
    """A kept token longer than length_t rejects; a long ignore run does not
    (backends/c/lexer.c:63-72)."""
    g = analysed["json"]  # length_t = u16
    assert oracle_rec(g, "lexer", b'"' + b"a" * 70000 + b'"').valid is False
    r = oracle_rec(g, "lexer", b"1" + b" " * 70000 + b"2")
    assert r.valid and r.terminals.tolist() == [0, 0]
    assert oracle_rec(g, "lexer", b"").valid
    assert oracle_rec(g, "lexer", b"   ").valid


# ------------------------------------------------------------ parser cross-checks

def token_alphabet(g):
    ids = [i for t, i in g.lex_both.term_id.items() if i != g.lex_both.ignore_id]
    return sorted(ids)


@pytest.mark.parametrize("name", ["arithmetic", "json", "sexp", "paper_grammar"])
def test_parser_vs_ll_exhaustive(analysed, name):
    g = analysed[name]
    alpha = token_alphabet(g)
    maxn = 4 if len(alpha) > 8 else 5
    for n in range(0, maxn + 1):
        for tup in itertools.product(alpha, repeat=n):
            ok, deriv = ll_parse(g, list(tup))
            r = oracle_rec(g, "parser", bytes(tup))
            assert r.valid == ok, tup
            if ok:
                assert r.productions.tolist() == deriv, tup


@pytest.mark.parametrize("name", ["arithmetic", "json", "sexp"])
def test_parser_vs_ll_random_derivations(analysed, name):
    g = analysed[name]
    rng = random.Random(42)
    prods = g.productions
    nterm = g.parser.end_t + 1
    start = prods[-1][1][1]
    by_lhs = {}
    for i, (lhs, rhs) in enumerate(prods):
        by_lhs.setdefault(lhs, []).append(rhs)
    for _ in range(60):
        form = [start]
        out = []
        budget = 400
        while form:
            s = form.pop(0)
            if s < nterm:
                out.append(s)
                continue
            alts = by_lhs[s]
            budget -= 1
            if budget < 0:  # steer towards short alternatives
                alts = sorted(alts, key=len)[:1] if len(out) > 50 else alts
            form = list(rng.choice(alts)) + form
            if len(form) > 200:
                break
        if form:
            continue
        for toks in (out, mutate(rng, out, token_alphabet(g))):
            ok, deriv = ll_parse(g, toks)
            r = oracle_rec(g, "parser", bytes(toks))
            assert r.valid == ok
            if ok:
                assert r.productions.tolist() == deriv


def mutate(rng, toks, alpha):
    t = list(toks)
    if t:
        t[rng.randrange(len(t))] = rng.choice(alpha)
    return t


@pytest.mark.parametrize("name", ["arithmetic", "json", "sexp"])
def test_parents_closed_form(analysed, name):
    """Combined records: parents of the compact tree equal the closed form
    applied to the full derivation, then renumbered."""
    from alpacc_amd import datagen
    try:
        datagen._load()
    except ImportError:
        pytest.skip("data generator not built")
    g = analysed[name]
    arity = g.parser.arity
    isterm = [t is not None for t in g.parser.prod_terminal]
    for seed in range(3):
        data = datagen.generate(name, 3000, seed, conforming=True).tobytes()
        r = oracle_rec(g, "both", data)
        assert r.valid, (name, seed)
        full = oracle_rec(g, "parser", r.terminals.tobytes())
        deriv = full.productions.tolist()
        par = closed_form_parents(deriv, arity)
        incl = np.cumsum([1 if isterm[p] else 0 for p in deriv])
        nodes = [p for p in deriv if not isterm[p]]
        tpar = [par[i] - incl[par[i]] for i, p in enumerate(deriv) if not isterm[p]]
        tokp = [par[i] - incl[par[i]] for i, p in enumerate(deriv) if isterm[p]]
        assert r.nodes.tolist() == nodes
        assert r.parents.tolist() == tpar
        assert r.token_parents.tolist() == tokp


# ---- width variants (--index32, length) ------------------------------------

WIDTH_VARIANTS = [(True, 0), (False, 8), (True, 64)]


def variant_grammar(name, index32, length):
    g = tables.load(os.path.join(ROOT, "grammars", name + ".alp"), index32=index32)
    if length:
        g.length_bits = length
    return g


@pytest.mark.parametrize("grammar", ["json", "arithmetic"])
@pytest.mark.parametrize("mode", ["lexer", "both"])
@pytest.mark.parametrize("index32,length", WIDTH_VARIANTS)
def test_width_variant_records(grammar, mode, index32, length):
    """A width variant's record is the default build's record with index
    and length fields re-encoded in the variant's widths, and a reject when
    a kept token does not fit the variant's length_t (the rule of
    backends/c/lexer.c:63-72 with length_t from Generator/C/Generator.hs:
    36-53): checked against the default oracle, which the known-answer tests
    above pin."""
    from alpacc_amd import datagen
    gd = variant_grammar(grammar, False, 0)
    gv = variant_grammar(grammar, index32, length)
    assert gv.index_bytes == (4 if index32 else 8)
    assert gv.length_bytes == (length // 8 if length else (gd.length_bits // 8 if gd.length_bits else gv.index_bytes))
    ld, lv = lex_layout(gd, mode), lex_layout(gv, mode)
    max_len = (1 << (8 * lv.length)) - 1
    inputs = [datagen.generate(grammar, n, seed=n, conforming=n % 2 == 0).tobytes() for n in (0, 1, 17, 300, 5000, 60000)]
    long_tok = b'"' + b"a" * 300 + b'"' if grammar == "json" else b"x" * 300 + b"+1"
    inputs += [long_tok, b"[" + long_tok + b"]" if grammar == "json" else b"(" + long_tok + b")"]
    for x in inputs:
        d = pyoracle.record(grammar, mode, x)
        v = pyoracle.record(grammar, mode, x, index32=index32, length=length)
        rec, end = wire.decode_record(d, 0, ld)
        assert end == len(d)
        if not rec.valid:
            # the default rejects: the same text with wider lengths may not,
            # but narrower ones never accept it
            if lv.length <= ld.length:
                assert v == b"\x00", (grammar, mode, len(x))
            continue
        exp = b"\x00" if len(rec.lengths) and int(rec.lengths.max()) > max_len else wire.encode_record(rec, lv)
        assert v == exp, (grammar, mode, index32, length, len(x))


def test_encode_record_inverts_decode():
    for grammar, mode in (("json", "lexer"), ("json", "both"), ("arithmetic", "both"), ("sexp", "parser")):
        g = variant_grammar(grammar, False, 0)
        lay = lex_layout(g, mode)
        payload = b"[1, 2, {}]" if grammar == "json" else (b"1+(2*a)" if grammar == "arithmetic" else bytes([0, 1]))
        rec_bytes = pyoracle.record(grammar, mode, payload)
        rec, _ = wire.decode_record(rec_bytes, 0, lay)
        assert wire.encode_record(rec, lay) == rec_bytes
