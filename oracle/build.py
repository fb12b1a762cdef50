"""Build recipe for the CPU oracle (test infrastructure, see oracle.c).

For every grammar under grammars/ and every build mode (lexer, parser, both)
this emits the constant block with alpacc_amd.gen.emit_c and compiles
oracle.c twice with gcc into oracle/build/:

  <grammar>_<mode>          the wire-protocol program (batch / --server)
  lib<grammar>_<mode>.so    the same code as an in-process library

Width variants (the generator's --index32 flag, a `length` other than the
grammar's) get a suffix: <grammar>_<mode>_i32, ..._l8, ..._i32_l8.

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


def variant_suffix(index32: bool = False, length: int = 0) -> str:
    return ("_i32" if index32 else "") + (f"_l{length}" if length else "")


def build_grammar(name: str, cc: str = "gcc", index32: bool = False, length: int = 0, modes=MODES,
                  programs: bool = True) -> None:
    g = tables.load(os.path.join(GRAMMARS, name + ".alp"), index32=index32)
    if length:
        g.length_bits = length
    sfx = variant_suffix(index32, length)
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(HERE, "oracle.c")
    for mode in modes:
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
    names = argv or sorted(os.path.splitext(f)[0] for f in os.listdir(GRAMMARS) if f.endswith(".alp"))
    for n in names:
        build_grammar(n)
        if n in VARIANT_GRAMMARS:
            for index32, length in VARIANTS:
                build_variant(n, index32, length)
        print(f"oracle: built {n}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
