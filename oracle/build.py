"""Build recipe for the CPU oracle (test infrastructure, see oracle.c).

For every grammar under grammars/ and every build mode (lexer, parser, both)
this emits the constant block with alpacc_amd.gen.emit_c and compiles
oracle.c twice with gcc into oracle/build/:

  <grammar>_<mode>          the wire-protocol program (batch / --server)
  lib<grammar>_<mode>.so    the same code as an in-process library

Width variants (the generator's --index32 flag, a `length` other than the
grammar's) get a suffix: <grammar>_<mode>_i32, ..._l8, ..._i32_l8.

Grammars come from grammars/ and (libraries only) from tests/golden/grammars/.

Usage: python oracle/build.py [grammar ...]
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

from alpacc_amd.gen import emit_c, tables  # noqa: E402

MODES = ("lexer", "parser", "both")
BUILD = os.path.join(HERE, "build")
GRAMMARS = os.path.join(ROOT, "grammars")
# grammars of the parity tests only (generated by alpacc_amd.gen random):
# oracles only, no programs
TEST_GRAMMARS = os.path.join(ROOT, "tests", "golden", "grammars")


def grammar_path(name: str) -> str:
    for d in (GRAMMARS, TEST_GRAMMARS):
        p = os.path.join(d, name + ".alp")
        if os.path.exists(p):
            return p
    raise FileNotFoundError(f"no grammar named {name}")


def grammar_names(d: str):
    if not os.path.isdir(d):
        return []
    return sorted(os.path.splitext(f)[0] for f in os.listdir(d) if f.endswith(".alp"))


def variant_suffix(index32: bool = False, length: int = 0) -> str:
    return ("_i32" if index32 else "") + (f"_l{length}" if length else "")


def build_grammar(name: str, cc: str = "gcc", index32: bool = False, length: int = 0, modes=MODES,
                  programs: bool = True) -> None:
    g = tables.load(grammar_path(name), index32=index32)
    if length:
        g.length_bits = length
    sfx = variant_suffix(index32, length)
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(HERE, "oracle.c")
    for mode in modes:
        if mode not in g.modes():  # a lexer-only grammar has no parser builds
            continue
        hdr = os.path.join(BUILD, f"{name}_{mode}{sfx}.h")
        text = emit_c.emit(g, mode)
        old = open(hdr).read() if os.path.exists(hdr) else None
        if old != text:
            with open(hdr, "w") as f:
                f.write(text)
        common = [cc, "-std=c99", "-D_POSIX_C_SOURCE=200809L", "-O2", "-Wall", "-Wno-unused-function", "-include", hdr, src]
        if programs:
            subprocess.check_call(common + ["-o", os.path.join(BUILD, f"{name}_{mode}{sfx}")])
        subprocess.check_call(common + ["-DORACLE_LIBRARY", "-fPIC", "-shared",
                                        "-o", os.path.join(BUILD, f"lib{name}_{mode}{sfx}.so")])


# width variants (libraries of the lexer and combined modes; the parser
# mode's records carry no index or length) checked by the parity tests
VARIANTS = ((True, 0), (False, 8), (True, 64))
VARIANT_GRAMMARS = ("json", "arithmetic")
VARIANT_MODES = ("lexer", "both")


def build_variant(name: str, index32: bool, length: int) -> None:
    build_grammar(name, index32=index32, length=length, modes=VARIANT_MODES, programs=False)


def main(argv):
    import concurrent.futures as cf
    names = argv or grammar_names(GRAMMARS)
    tests = [] if argv else grammar_names(TEST_GRAMMARS)
    for n in names:
        build_grammar(n)
        if n in VARIANT_GRAMMARS:
            for index32, length in VARIANTS:
                build_variant(n, index32, length)
        print(f"oracle: built {n}", flush=True)
    if tests:
        with cf.ThreadPoolExecutor(max_workers=8) as ex:
            list(ex.map(lambda n: build_grammar(n, programs=False), tests))
        print(f"oracle: built {len(tests)} test grammars", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
