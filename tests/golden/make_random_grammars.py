"""Writes the random grammars of the parity tests (tests/golden/grammars/).

The reference's differential suite draws grammars with `alpacc random` and
runs them for (q, k) in {(0, 0) lexer-only, (0, 1), (1, 1), (2, 2), (3, 3)}
(Makefile:80-88, tests/testc.sh:60-144), skipping those the generator
refuses.  Here the same kinds of grammars come from alpacc_amd.gen.randgrammar
with fixed seeds: for each (q, k), the first PER_SETTING seeds whose grammar
the table generator accepts (LLP(q, k), within the device tables' limits).
The (0, 0) grammars have terminal rules only (lexer mode).

Usage: python tests/golden/make_random_grammars.py
"""

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from alpacc_amd.gen import emit_hip, tables  # noqa: E402
from alpacc_amd.gen.grammar import GrammarError  # noqa: E402
from alpacc_amd.gen.randgrammar import random_grammar  # noqa: E402

SETTINGS = [(0, 0), (0, 1), (1, 1), (2, 2), (3, 3)]
PER_SETTING = 5
OUT = os.path.join(HERE, "grammars")


def accepted(name, text):
    try:
        g = tables.analyse(name, text)
        emit_hip.device_tables(g)
        return True
    except (GrammarError, emit_hip.EmitError, ValueError):
        return False


# Grammars of a chosen shape, for the widest cases of the kernels: windows
# with more than four productions (the resolve kernel's widest
# instantiation), bracket symbols beyond 255 (two-byte bracket stacks), 268
# productions (two-byte production ids), and a lexer of 26 keywords beside identifiers whose byte-step table (1300
# endomorphisms x 30 byte classes, 78,000 bytes) is too large for the fold's
# 16-bit row offsets, so it takes the fold kernel with plain ids.
SHAPED = {
    "wide_windows": """params {
  lookback = 1.
  lookahead = 1.
}

x = /x/.
y = /y/.
z = /z/.
ignore = /\\s+/.

S -> "[" L "]".
L -> | I L.
I -> A | W.
A -> B.
B -> C.
C -> D.
D -> x.
W -> y S S z.
""",
    "many_symbols": "\n".join(
        ["params {", "  lookback = 1.", "  lookahead = 1.", "}", "", "c = /c/.", "ignore = /\\s+/.", "",
         "S -> N0."]
        + [f'N{i} -> "a{i}" {f"N{i + 1}" if i + 1 < 60 else "c"} "b{i}" | c.' for i in range(60)]) + "\n",
    "many_productions": "\n".join(
        ["params {", "  lookback = 1.", "  lookahead = 1.", "}", "", "c = /c/.", "ignore = /\\s+/.", "",
         "S -> N0."]
        + [f'N{i} -> "a{i}" {f"N{i + 1}" if i + 1 < 66 else "c"} "b{i}" | c.' for i in range(66)]) + "\n",
    "keywords": "".join(
        f"k{w} = /{w}/.\n"
        for w in ["if", "then", "else", "while", "do", "end", "return", "for", "function", "switch", "case",
                  "default", "break", "continue", "struct", "static", "signed", "unsigned", "typedef", "union",
                  "volatile", "register", "extern", "goto", "sizeof", "inline"])
    + 'id = /[a-z]+[0-9]*/.\nnum = /-?[0-9]+(.[0-9]+)?/.\nstr = /"[a-z ]*"/.\nignore = /\\s+/.\n',
}


def main():
    os.makedirs(OUT, exist_ok=True)
    for name, text in SHAPED.items():
        assert accepted(name, text), name
        with open(os.path.join(OUT, name + ".alp"), "w") as f:
            f.write(f"# grammar of a chosen shape (tests/golden/make_random_grammars.py)\n" + text)
    for q, k in SETTINGS:
        found = 0
        seed = 0
        while found < PER_SETTING:
            seed += 1
            lexer_only = (q, k) == (0, 0)
            text = random_grammar(1000 * q + 100 * k + seed, num_productions=0 if lexer_only else 3,
                                  q=None if lexer_only else q, k=None if lexer_only else k)
            name = f"rand_q{q}k{k}_s{seed}"
            if accepted(name, text):
                with open(os.path.join(OUT, name + ".alp"), "w") as f:
                    f.write(f"# random grammar: alpacc_amd.gen.randgrammar seed {1000 * q + 100 * k + seed}\n" + text)
                found += 1
        print(f"q={q} k={k}: {found} grammars (seeds 1..{seed})")


if __name__ == "__main__":
    main()
