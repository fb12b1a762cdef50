// Device-table descriptor of one grammar, filled by the sources that
// alpacc_amd/gen/emit_hip.py generates (one per grammar).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace alpacc {

struct GrammarDesc {
  const char* name;
  int q, k;
  // ---- lexer ----
  int nstate;                    // DFA states including the dead state 0; 1 is initial
  int nendo;                     // endomorphisms; 0 = identity, 1 = dead
  int nclass;                    // byte classes
  const uint8_t* byte_class;     // [256]
  const uint16_t* step;          // [nendo * nclass]  e then class c
  const uint16_t* compose;       // [nendo * nendo]   a then b
  const uint8_t* endo_init;      // [nendo]           state the initial state maps to
  const uint16_t* delta_lex;     // [nstate * 256]    next | produce<<8 | kept<<9
  const uint16_t* delta_both;    // same (kept: the token state s accepts is not ignored)
  const uint8_t* token_lex;      // [nstate]          terminal state s accepts
  const uint8_t* token_both;     // same, combined-mode terminal ids
  int ignore_lex, ignore_both;   // -1 when the grammar has no ignore terminal
  int dead_lex, dead_both;       // terminal id of "no token"
  int terminal_bytes_lex, terminal_bytes_both, index_bytes, length_bytes;
  // ---- LLP parser ----
  int empty_t, start_t, end_t;
  int window_base;               // terminal symbols (window digits are base-window_base)
  int nwindow;
  const uint16_t* window;        // [nwindow] pair id or 0xffff
  int npairs;
  const int32_t* spans;          // [npairs * 4] stack start, stack len, prod start, prod len
  const uint8_t* profile;        // [npairs * 8] i8 ddepth, i8 dmin, i8 dslot, i8 smin, u8 terminals, pad
  int nstacks;
  const uint16_t* stacks;        // bracket: symbol | left << 15
  int nprods_flat;
  const uint16_t* productions;   // production strings
  int nproductions;
  int production_bytes;
  const uint8_t* arity;          // [nproductions]
  const uint8_t* prod_is_terminal;  // [nproductions]
  int max_brackets, max_prods;   // per window
  int max_slot_items;            // per window: productions + child slots they open
  const uint8_t* pair_rec;       // [(npairs + 1) * 64] window net effects (layout in emit_hip.py);
                                 // record npairs is all zero
  const uint8_t* prod_info;      // [nproductions] arity | terminal << 7
  int max_bpop, max_bpush;       // per window, after cancellation inside the window
  int max_spop, max_spush;
  int max_nt;                    // terminal productions per window
  int length_param;              // the grammar's `length` (bits), 0 when it has none: length_t follows index_t
  int has_parser;                // 0: terminal rules only, a lexer-only grammar (its parser fields are empty)
  // names by id, NUL-separated: terminals in lexer-only and in combined
  // numbering (the dead terminal last), productions
  const uint8_t* terminal_names_lex;
  const uint8_t* terminal_names_both;
  const uint8_t* production_names;
  // 0: window is a dense index by window code (nwindow = window_base^(q+k)
  // entries); else log2 of the slots of an open-addressing table of the
  // windows with a pair: window_keys[slot] the window code (all ones when
  // empty) and window[slot] its pair id, nwindow slots, first slot = the
  // code times 0x9E3779B97F4A7C15, top window_hash bits, then linear probing
  int window_hash;
  const uint64_t* window_keys;   // [nwindow] when hashed, else one unused entry
  // production_bytes 2 (more than 255 productions): the productions of
  // every pair record as u16 ids, 8 a pair (the records hold their low
  // bytes); else one unused entry
  const uint16_t* pair_prod16;
};

// A grammar by name: one loaded at run time (alpacc_hip_load_grammar), else
// one compiled into the library.
const GrammarDesc* find_grammar(const char* name);
// The compiled-in grammars (generated registry).
const GrammarDesc* find_builtin_grammar(const char* name);
// Names of every grammar, compiled-in and loaded, null-terminated.
const char* const* grammar_names();
const char* const* builtin_grammar_names();
// Parses a tables file (alpacc_amd/gen/emit_hip.py emit_tables) into a
// descriptor whose arrays live in *storage; returns 0 or an error.
int parse_grammar_tables(const uint8_t* data, size_t len, GrammarDesc* out, std::vector<std::vector<uint8_t>>* storage,
                         std::string* name);

}  // namespace alpacc
