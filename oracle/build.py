"""Build recipe for the CPU oracle (test infrastructure, see oracle.c).

For every grammar under grammars/ and every build mode (lexer, parser, both)
this emits the constant block with alpacc_amd.gen.emit_c and compiles
oracle.c twice with gcc into oracle/build/:

  <grammar>_<mode>          the wire-protocol program (batch / --server)
  lib<grammar>_<mode>.so    the same code as an in-process library

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


def build_grammar(name: str, cc: str = "gcc") -> None:
    g = tables.load(os.path.join(GRAMMARS, name + ".alp"))
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(HERE, "oracle.c")
    for mode in MODES:
        hdr = os.path.join(BUILD, f"{name}_{mode}.h")
        text = emit_c.emit(g, mode)
        old = open(hdr).read() if os.path.exists(hdr) else None
        if old != text:
            with open(hdr, "w") as f:
                f.write(text)
        common = [cc, "-std=c99", "-D_POSIX_C_SOURCE=200809L", "-O2", "-Wall", "-Wno-unused-function", "-include", hdr, src]
        subprocess.check_call(common + ["-o", os.path.join(BUILD, f"{name}_{mode}")])
        subprocess.check_call(common + ["-DORACLE_LIBRARY", "-fPIC", "-shared",
                                        "-o", os.path.join(BUILD, f"lib{name}_{mode}.so")])


def main(argv):
    names = argv or sorted(os.path.splitext(f)[0] for f in os.listdir(GRAMMARS) if f.endswith(".alp"))
    for n in names:
        build_grammar(n)
        print(f"oracle: built {n}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
