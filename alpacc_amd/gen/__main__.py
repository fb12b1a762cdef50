"""The generator's command line: the `alpacc` subcommands of this target
(app/Main.hs:393-402).

  python -m alpacc_amd.gen hip GRAMMAR.alp [-o FILE] [--source FILE.cpp]
      the `alpacc hip` target: the grammar's device tables as a tables file
      for alpacc_hip_load_grammar / `alpacc-hip --tables FILE` (default
      GRAMMAR.alphip), and optionally as the C++ source that compiles them
      into the library (what alpacc_amd/build.py does for grammars/);
  python -m alpacc_amd.gen random [-o FILE] [--seed N] [--num-chars 3]
      [--num-terminals 3] [--num-nonterminals 2] [--num-productions 3]
      [--lookback Q --lookahead K]
      a random grammar (`alpacc random`);
  python -m alpacc_amd.gen names GRAMMAR.alp [--lexer]
      the id -> name tables (the Futhark target's terminal / production types);
  python -m alpacc_amd.gen dev {composition-histogram,image-sizes,
      raw-image-sizes,dump-endo} GRAMMAR.alp [CHAR]
      lexer diagnostics (`alpacc dev`).
"""

from __future__ import annotations

import argparse
import os
import sys

from . import dev, emit_hip, tables
from .randgrammar import random_grammar


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m alpacc_amd.gen")
    sub = ap.add_subparsers(dest="cmd", required=True)
    h = sub.add_parser("hip", help="device tables of a grammar")
    h.add_argument("grammar")
    h.add_argument("-o", "--output")
    h.add_argument("--source", help="also write the C++ source of the tables")
    r = sub.add_parser("random", help="a random grammar")
    r.add_argument("-o", "--output")
    r.add_argument("--seed", type=int, default=0)
    r.add_argument("--num-chars", type=int, default=3)
    r.add_argument("--num-terminals", type=int, default=3)
    r.add_argument("--num-nonterminals", type=int, default=2)
    r.add_argument("--num-productions", type=int, default=3)
    r.add_argument("--lookback", type=int)
    r.add_argument("--lookahead", type=int)
    n = sub.add_parser("names", help="id -> name tables")
    n.add_argument("grammar")
    n.add_argument("--lexer", action="store_true")
    d = sub.add_parser("dev", help="lexer diagnostics")
    d.add_argument("what", choices=["composition-histogram", "image-sizes", "raw-image-sizes", "dump-endo"])
    d.add_argument("grammar")
    d.add_argument("char", nargs="?", type=int)
    a = ap.parse_args(argv)

    if a.cmd == "random":
        text = random_grammar(a.seed, a.num_chars, a.num_terminals, a.num_nonterminals, a.num_productions,
                              q=a.lookback, k=a.lookahead)
        if a.output:
            with open(a.output, "w") as f:
                f.write(text)
        else:
            sys.stdout.write(text)
        return 0
    try:
        g = tables.load(a.grammar)
    except Exception as e:  # the reference prints the generator's error and fails
        print(f"error: {e}", file=sys.stderr)
        return 1
    if a.cmd == "hip":
        out = a.output or os.path.splitext(os.path.basename(a.grammar))[0] + ".alphip"
        try:
            blob = emit_hip.emit_tables(g)
            src = emit_hip.emit(g) if a.source else None
        except emit_hip.EmitError as e:
            print(f"error: {e}", file=sys.stderr)
            return 1
        with open(out, "wb") as f:
            f.write(blob)
        if a.source:
            with open(a.source, "w") as f:
                f.write(src)
        return 0
    if a.cmd == "names":
        lt = g.lex_only if a.lexer else g.lex_both
        for i, nm in enumerate(lt.names()):
            print(f"terminal {i} {nm}")
        if not a.lexer and g.has_parser:
            for i, nm in enumerate(g.parser.prod_names):
                print(f"production {i} {nm}")
        return 0
    if a.what == "composition-histogram":
        sys.stdout.write(dev.composition_histogram(g))
    elif a.what == "dump-endo":
        if a.char is None:
            print("error: dump-endo needs a byte value", file=sys.stderr)
            return 1
        sys.stdout.write(dev.dump_endo(g, a.char))
    else:
        sys.stdout.write(dev.image_sizes(g, raw=a.what == "raw-image-sizes"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
