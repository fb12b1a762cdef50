"""Emit the per-grammar constant block consumed by the C oracle.

The block has the shape the reference C generator writes before its runtime
templates (src/Alpacc/Generator/C/Lexer.hs:16-58, C/Parser.hs:21-95,
C/Generator.hs:36-53): typedefs, masks, TO_STATE / COMPOSE / ACCEPT for the
lexer and STACKS / PRODUCTIONS / HASH_TABLE_* for the LLP parser, plus the
tree extras for the combined mode.
"""

from __future__ import annotations

from typing import List

from .tables import Grammar

_CT = {1: "uint8_t", 2: "uint16_t", 4: "uint32_t", 8: "uint64_t"}


def _arr(vals, per_line=16) -> str:
    vals = [str(v) for v in vals]
    if not vals:
        return "{0}"
    lines = []
    for i in range(0, len(vals), per_line):
        lines.append("  " + ", ".join(vals[i:i + per_line]))
    return "{\n" + ",\n".join(lines) + "\n}"


def _b(v: bool) -> str:
    return "true" if v else "false"


def emit(g: Grammar, mode: str) -> str:
    """mode: 'lexer' | 'parser' | 'both'."""
    assert mode in ("lexer", "parser", "both")
    if mode != "lexer" and not g.has_parser:
        raise ValueError(f"{g.name}: CFG has no production rules (lexer mode only)")
    out: List[str] = [f"/* generated by alpacc_amd.gen for grammar '{g.name}', mode {mode} */",
                      "#include <stdbool.h>", "#include <stdint.h>"]
    tbytes = g.terminal_bytes_lex if mode == "lexer" else g.terminal_bytes_both
    out.append(f"typedef {_CT[tbytes]} terminal_t;")
    out.append("typedef int32_t index_t;" if g.index32 else "typedef int64_t index_t;")
    if mode != "parser":
        out.append(f"typedef {_CT[g.length_bytes]} length_t;")
        out.append("#define HAS_LEXER 1")
    if mode != "lexer":
        out.append("#define HAS_PARSER 1")
    if mode == "both":
        out.append("#define WITH_TREE 1")
    if mode != "parser":
        lt = g.lex_only if mode == "lexer" else g.lex_both
        N = lt.endos.size
        eb, tb = lt.endo_bits, lt.term_bits
        out += [
            f"typedef {_CT[lt.state_bytes]} state_t;",
            f"#define NUM_STATES {N}",
            f"#define ENDO_MASK ((state_t) {(1 << eb) - 1})",
            "#define ENDO_OFFSET 0",
            f"#define TERMINAL_MASK ((state_t) {((1 << tb) - 1) << eb})",
            f"#define TERMINAL_OFFSET {eb}",
            f"#define PRODUCE_MASK ((state_t) {1 << (eb + tb)})",
            f"#define PRODUCE_OFFSET {eb + tb}",
            f"#define IDENTITY ((state_t) {lt.identity})",
        ]
        if lt.ignore_id is not None:
            out.append(f"#define IGNORE_TOKEN ((terminal_t) {lt.ignore_id})")
        out.append(f"static const state_t TO_STATE[256] = {_arr(lt.to_state)};")
        out.append(f"static const state_t COMPOSE[NUM_STATES * NUM_STATES] = {_arr(lt.compose, 24)};")
        out.append(f"static const bool ACCEPT[NUM_STATES] = {_arr([_b(a) for a in lt.accept])};")
    if mode != "lexer":
        p = g.parser
        out += [
            f"typedef {_CT[p.production_bytes]} production_t;",
            f"typedef {_CT[p.bracket_bytes]} bracket_t;",
            f"#define Q {p.q}",
            f"#define K {p.k}",
            f"#define EMPTY_TERMINAL ((terminal_t) {p.empty_t})",
            f"#define START_TERMINAL ((terminal_t) {p.start_t})",
            f"#define END_TERMINAL ((terminal_t) {p.end_t})",
            f"#define HASH_TABLE_SIZE {p.hash_size}",
            f"#define MAX_ITERS {p.max_iters}",
            f"#define STACKS_SIZE {max(1, len(p.stacks))}",
            f"#define PRODUCTIONS_SIZE {max(1, len(p.productions))}",
            f"#define MAX_BRACKETS_PER_POSITION {max(1, p.max_brackets)}",
            f"#define MAX_PRODS_PER_POSITION {max(1, p.max_prods)}",
            f"static const bracket_t STACKS[STACKS_SIZE] = {_arr(p.stacks)};",
            f"static const production_t PRODUCTIONS[PRODUCTIONS_SIZE] = {_arr(p.productions)};",
            f"static const bool HASH_TABLE_IS_VALID[HASH_TABLE_SIZE] = {_arr([_b(v) for v in p.hash_valid])};",
            "static const terminal_t HASH_TABLE_KEYS[HASH_TABLE_SIZE][Q + K] = "
            + _arr(["{" + ", ".join(str(x) for x in key) + "}" for key in p.hash_keys], 8) + ";",
            "static const int64_t HASH_TABLE_STACKS_SPAN[HASH_TABLE_SIZE][2] = "
            + _arr([f"{{{s[0]}, {s[1]}}}" for s in p.hash_spans], 8) + ";",
            "static const int64_t HASH_TABLE_PRODUCTIONS_SPAN[HASH_TABLE_SIZE][2] = "
            + _arr([f"{{{s[2]}, {s[3]}}}" for s in p.hash_spans], 8) + ";",
        ]
        if mode == "both":
            out += [
                f"#define NUMBER_OF_PRODUCTIONS {p.n_productions}",
                f"static const uint64_t PRODUCTION_TO_ARITY[NUMBER_OF_PRODUCTIONS] = {_arr(p.arity)};",
                "static const bool PRODUCTION_TO_TERMINAL_IS_VALID[NUMBER_OF_PRODUCTIONS] = "
                + _arr([_b(t is not None) for t in p.prod_terminal]) + ";",
            ]
    return "\n".join(out) + "\n"
