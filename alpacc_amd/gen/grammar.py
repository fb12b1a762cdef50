"""Reader for alpacc grammar files (``.alp``) and their regular expressions.

This is the front end of the ``alpacc hip`` table generator.  It follows the
grammar-file syntax accepted by the reference front end:

* ``params { lookback = N. lookahead = N. length = 8|16|32|64. }`` (optional,
  defaults q=0, k=1, no length) -- src/Alpacc/CFG.hs:51-103
* terminal rules ``name = /regex/.`` -- src/Alpacc/CFG.hs:202-209
* nonterminal rules ``NT [Name] -> rhs | rhs .`` -- src/Alpacc/CFG.hs:211-223
* whitespace and ``#`` line comments between tokens -- src/Alpacc/CFG.hs:165-167

and the regex dialect of src/Alpacc/Lexer/RegularExpression.hs:81-220
(``.`` is a literal dot, ``\\s`` is one space byte, whitespace inside ``/../``
is skipped, ``R+`` = ``R R*``, ``R?`` = ``(|R)``).

Regexes are kept as small tuples over *byte strings* (every character is its
UTF-8 encoding, RegularExpression.hs:64-65):

    ('eps',) | ('lit', bytes) | ('range', [bytes, ...]) | ('star', r)
    | ('alt', r, r) | ('cat', r, r)
"""

from __future__ import annotations

import dataclasses
import unicodedata
from typing import List, Optional, Tuple


class GrammarError(ValueError):
    pass


# --------------------------------------------------------------------------
# Regular expressions (src/Alpacc/Lexer/RegularExpression.hs)
# --------------------------------------------------------------------------

# RegularExpression.hs:84-99 -- the characters a backslash may escape.
_ESCAPES = {
    "\\": "\\", "/": "/", "[": "[", "]": "]", "(": "(", ")": ")",
    "*": "*", "+": "+", "|": "|", "s": " ", "t": "\t", "n": "\n",
    "r": "\r", "?": "?",
}
_SPECIAL = set("\\/[]()*|+?")


def _is_print(c: str) -> bool:
    # Haskell isPrint: not a control, separator-other, surrogate, private or
    # unassigned code point.
    cat = unicodedata.category(c)
    if c == " ":
        return True
    return not (cat.startswith("C") or cat in ("Zl", "Zp"))


def _is_space(c: str) -> bool:
    return c.isspace()


def _utf8(c: str) -> bytes:
    return c.encode("utf-8")


class _RegexParser:
    """Recursive descent over the regex text (RegularExpression.hs:172-220)."""

    def __init__(self, text: str, pos: int):
        self.s = text
        self.i = pos

    # ``lexeme``: skip whitespace after every token (space1, no comments).
    def _ws(self):
        while self.i < len(self.s) and _is_space(self.s[self.i]):
            self.i += 1

    def _peek(self) -> str:
        return self.s[self.i] if self.i < len(self.s) else ""

    def parse_alter(self):
        r = self.parse_concat()
        while self._peek() == "|":
            self.i += 1
            self._ws()
            r = ("alt", r, self.parse_concat())
        return r

    def parse_concat(self):
        # foldl Concat Epsilon <$> many pTerm (RegularExpression.hs:166-167)
        r = ("eps",)
        while True:
            t = self.try_term()
            if t is None:
                return r
            r = ("cat", r, t)

    def try_term(self):
        c = self._peek()
        if c == "[":
            term = self.parse_range()
        elif c == "(":
            self.i += 1
            self._ws()
            term = self.parse_alter()
            if self._peek() != ")":
                raise GrammarError(f"expected ')' at offset {self.i}")
            self.i += 1
            self._ws()
        else:
            lit = self.try_char()
            if lit is None:
                return None
            term = ("lit", lit)
        # postfix operators (RegularExpression.hs:205-212)
        while True:
            c = self._peek()
            if c == "*":
                term = ("star", term)
            elif c == "+":
                term = ("cat", term, ("star", term))
            elif c == "?":
                term = ("alt", ("eps",), term)
            else:
                return term
            self.i += 1
            self._ws()

    def _hex2(self, at: int) -> Optional[int]:
        h = self.s[at:at + 2]
        if len(h) == 2 and all(ch in "0123456789abcdefABCDEF" for ch in h):
            return int(h, 16)
        return None

    def try_char(self) -> Optional[bytes]:
        """pChar: escape | \\u hex+ | \\x hh | printable non-special char."""
        s, i = self.s, self.i
        if i >= len(s):
            return None
        c = s[i]
        if c == "\\":
            nxt = s[i + 1:i + 2]
            if nxt in _ESCAPES:
                self.i = i + 2
                self._ws()
                return _utf8(_ESCAPES[nxt])
            if nxt == "u":
                j = i + 2
                while j < len(s) and s[j] in "0123456789abcdefABCDEF":
                    j += 1
                if j > i + 2:
                    self.i = j
                    self._ws()
                    return _utf8(chr(int(s[i + 2:j], 16)))
            if nxt == "x":
                v = self._hex2(i + 2)
                if v is not None:
                    self.i = i + 4
                    self._ws()
                    return bytes([v])
            raise GrammarError(f"bad escape at offset {i}")
        if c in _SPECIAL or not _is_print(c) or _is_space(c):
            return None
        self.i = i + 1
        self._ws()
        return _utf8(c)

    def parse_range(self):
        # RegularExpression.hs:180-199
        assert self._peek() == "["
        self.i += 1
        self._ws()
        items: List[bytes] = []
        s = self.s
        while True:
            c = self._peek()
            if c == "]":
                break
            if c == "\\" and s[self.i + 1:self.i + 2] == "x":
                a = self._hex2(self.i + 2)
                if a is None:
                    raise GrammarError("bad \\x in range")
                self.i += 4
                self._ws()
                if self._peek() == "-" and s[self.i + 1:self.i + 3] == "\\x":
                    save = self.i
                    self.i += 1
                    self._ws()
                    b = self._hex2(self.i + 2)
                    if b is not None:
                        self.i += 4
                        self._ws()
                        items.extend(bytes([v]) for v in range(a, b + 1))
                        continue
                    self.i = save
                items.append(bytes([a]))
                continue
            if c and c.isalnum():
                self.i += 1
                self._ws()
                if self._peek() == "-":
                    save = self.i
                    self.i += 1
                    self._ws()
                    d = self._peek()
                    if d and d.isalnum():
                        self.i += 1
                        self._ws()
                        items.extend(_utf8(chr(v)) for v in range(ord(c), ord(d) + 1))
                        continue
                    self.i = save
                items.append(_utf8(c))
                continue
            raise GrammarError(f"bad range item at offset {self.i}")
        if not items:
            raise GrammarError("empty range")
        self.i += 1
        self._ws()
        return ("range", items)


def parse_regex(text: str):
    p = _RegexParser(text, 0)
    p._ws()
    r = p.parse_alter()
    if p.i != len(text):
        raise GrammarError(f"trailing regex text at offset {p.i}")
    return r


def produces_epsilon(r) -> bool:
    """RegularExpression.hs:53-59."""
    tag = r[0]
    if tag == "eps" or tag == "star":
        return True
    if tag in ("lit", "range"):
        return False
    if tag == "alt":
        return produces_epsilon(r[1]) or produces_epsilon(r[2])
    return produces_epsilon(r[1]) and produces_epsilon(r[2])


# --------------------------------------------------------------------------
# Grammar files (src/Alpacc/CFG.hs)
# --------------------------------------------------------------------------

# A terminal is ('T', name) for a named terminal or ('L', text) for a string
# literal used in a production.  Python orders these tuples the way Haskell's
# derived Ord orders ``T Text | TLit Text``: every named terminal before every
# literal, then by text (code points).


@dataclasses.dataclass
class Params:
    lookback: int = 0
    lookahead: int = 1
    length: Optional[int] = None


@dataclasses.dataclass
class TRule:
    term: Tuple[str, str]
    regex: tuple


@dataclasses.dataclass
class NTRule:
    nt: str
    name: Optional[str]
    rhs: List[Tuple[str, object]]  # ('T', term) | ('N', name)


@dataclasses.dataclass
class CFG:
    params: Params
    t_rules: List[TRule]
    nt_rules: List[NTRule]


class _CFGParser:
    def __init__(self, text: str):
        self.s = text
        self.i = 0

    def ws(self):
        s = self.s
        while self.i < len(s):
            c = s[self.i]
            if c.isspace():
                self.i += 1
            elif c == "#":
                while self.i < len(s) and s[self.i] != "\n":
                    self.i += 1
            else:
                break

    def peek(self, n=1):
        return self.s[self.i:self.i + n]

    def expect(self, tok: str):
        if not self.s.startswith(tok, self.i):
            raise GrammarError(f"expected {tok!r} at offset {self.i}")
        self.i += len(tok)
        self.ws()

    def accept(self, tok: str) -> bool:
        if self.s.startswith(tok, self.i):
            self.i += len(tok)
            self.ws()
            return True
        return False

    def ident(self, first_ok) -> Optional[str]:
        s, i = self.s, self.i
        if i < len(s) and first_ok(s[i]):
            j = i + 1
            while j < len(s) and (s[j] == "_" or (s[j].isascii() and s[j].isalnum())):
                j += 1
            self.i = j
            self.ws()
            return s[i:j]
        return None

    def decimal(self) -> int:
        s, i = self.s, self.i
        j = i
        while j < len(s) and s[j].isdigit():
            j += 1
        if j == i:
            raise GrammarError(f"expected a number at offset {i}")
        self.i = j
        self.ws()
        return int(s[i:j])

    def params(self) -> Params:
        p = Params()
        save = self.i
        if not self.accept("params"):
            return p
        if not self.accept("{"):
            self.i = save
            return p
        while not self.accept("}"):
            for key in ("lookback", "lookahead", "length"):
                if self.s.startswith(key, self.i):
                    self.i += len(key)
                    self.ws()
                    self.expect("=")
                    n = self.decimal()
                    self.expect(".")
                    if key == "lookback":
                        p.lookback = n
                    elif key == "lookahead":
                        p.lookahead = n
                    else:
                        if n not in (8, 16, 32, 64):
                            raise GrammarError("length must be 8, 16, 32, or 64")
                        p.length = n
                    break
            else:
                raise GrammarError(f"bad parameter at offset {self.i}")
        return p

    def t_rule(self) -> Optional[TRule]:
        save = self.i
        name = self.ident(lambda c: "a" <= c <= "z")
        if name is None or not self.accept("="):
            self.i = save
            return None
        if self.peek() != "/":
            self.i = save
            return None
        self.i += 1
        self.ws()
        rp = _RegexParser(self.s, self.i)
        rp._ws()
        regex = rp.parse_alter()
        self.i = rp.i
        if self.peek() != "/":
            raise GrammarError(f"unterminated regex at offset {self.i}")
        self.i += 1
        self.ws()
        self.expect(".")
        return TRule(("T", name), regex)

    def string_lit(self) -> Optional[str]:
        s, i = self.s, self.i
        if s[i:i + 1] != '"':
            return None
        j = i + 1
        while j < len(s) and s[j] != '"' and _is_print(s[j]):
            j += 1
        if j == i + 1 or s[j:j + 1] != '"':
            raise GrammarError(f"bad string literal at offset {i}")
        self.i = j + 1
        self.ws()
        return s[i + 1:j]

    def symbol(self):
        t = self.ident(lambda c: "a" <= c <= "z")
        if t is not None:
            return ("T", ("T", t))
        lit = self.string_lit()
        if lit is not None:
            return ("T", ("L", lit))
        nt = self.ident(lambda c: "A" <= c <= "Z")
        if nt is not None:
            return ("N", nt)
        return None

    def nt_rules(self) -> Optional[List[NTRule]]:
        nt = self.ident(lambda c: "A" <= c <= "Z")
        if nt is None:
            return None
        name = None
        if self.accept("["):
            name = self.ident(lambda c: "A" <= c <= "Z")
            if name is None:
                raise GrammarError("production name must start uppercase")
            self.expect("]")
        self.expect("->")
        rules = []
        while True:
            rhs = []
            while True:
                sym = self.symbol()
                if sym is None:
                    break
                rhs.append(sym)
            rules.append(NTRule(nt, name, rhs))
            if not self.accept("|"):
                break
        self.expect(".")
        return rules

    def cfg(self) -> CFG:
        self.ws()
        params = self.params()
        t_rules = []
        while True:
            r = self.t_rule()
            if r is None:
                break
            t_rules.append(r)
        nt_rules: List[NTRule] = []
        while True:
            rs = self.nt_rules()
            if rs is None:
                break
            nt_rules.extend(rs)
        if self.i != len(self.s):
            raise GrammarError(f"unexpected text at offset {self.i}")
        return CFG(params, t_rules, nt_rules)


def parse_cfg(text: str) -> CFG:
    return _CFGParser(text).cfg()


def literal_regex(text: str):
    """A production literal becomes the concatenation of its characters
    (CFG.hs:145-146)."""
    chars = [("lit", _utf8(c)) for c in text]
    r = chars[0]
    for c in chars[1:]:
        r = ("cat", r, c)
    return r
