/*
 * alpacc_hip.h -- C ABI of the MI355X (gfx950) backend for alpacc lexers and
 * LLP parsers.
 *
 * The reference has no shared library: its boundary is a set of file-static
 * functions in the generated C program plus a byte-level wire protocol.  Each
 * entry point below names the reference function it replaces.  All pointers
 * are plain host or device pointers; no HIP or torch types appear here (a
 * stream is passed as an opaque pointer, NULL = the context's own stream).
 *
 * Return codes: 1 = input accepted, 0 = input rejected (the wire protocol's
 * 0x00 record), negative = error (alpacc_hip_last_error() has the message).
 *
 * Build modes (lexer-only, parser-only, combined) are a runtime argument; the
 * terminal numbering of each mode is that of the reference's build of the
 * same mode (lexer-only numbers terminals in sorted order).
 */
#ifndef ALPACC_HIP_H
#define ALPACC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct alpacc_hip_ctx alpacc_hip_ctx;

enum alpacc_hip_mode { ALPACC_MODE_LEXER = 0, ALPACC_MODE_PARSER = 1, ALPACC_MODE_BOTH = 2 };

enum alpacc_hip_status {
  ALPACC_VALID = 1,
  ALPACC_REJECT = 0,
  ALPACC_ERR_ARG = -1,     /* bad argument or unknown grammar */
  ALPACC_ERR_HIP = -2,     /* HIP runtime failure */
  ALPACC_ERR_NOMEM = -3,   /* device or host allocation failed */
  ALPACC_ERR_LIMIT = -4    /* input beyond a documented size limit */
};

/* Tokens as structure-of-arrays: terminal_t, index_t (int64) and length_t
 * arrays of `count` entries each.  Same content as the reference's lexeme_t
 * array (backends/c/lexer.c:30-34). */
typedef struct {
  uint64_t count;
  void *terminals;
  int64_t *starts;
  void *lengths;
} alpacc_hip_lexemes;

/* Combined-mode result: the tokens plus the compact productions-only tree of
 * docs/wire-protocols.md:108-125 (preorder production ids, parent indices,
 * and each token's parent node). */
typedef struct {
  alpacc_hip_lexemes tokens;
  uint64_t num_nodes;
  void *productions;       /* production_t[num_nodes] */
  int64_t *parents;        /* index_t[num_nodes] */
  int64_t *token_parents;  /* index_t[tokens.count] */
} alpacc_hip_tree;

/* ---- context ------------------------------------------------------------ */

/* Make one grammar's tables resident on `device` (replaces the generated
 * program's static const tables, src/Alpacc/Generator/C/Lexer.hs:16-58 and
 * C/Parser.hs:21-95). */
int alpacc_hip_open(const char *grammar, int device, alpacc_hip_ctx **out);

/* A context whose integer widths follow the generator's --index32 flag and
 * the grammar's `length` parameter (src/Alpacc/Generator/C/Generator.hs:
 * 36-53, docs/wire-protocols.md:37-50) chosen at open time instead of at
 * generation time: index32 != 0 makes index_t int32 (starts, parents);
 * length_bits (8, 16, 32, 64) sets length_t, 0 keeps the grammar's `length`,
 * or without one the unsigned counterpart of index_t.  Records, layouts and
 * the length overflow rule follow these widths. */
int alpacc_hip_open_variant(const char *grammar, int device, int index32, int length_bits, alpacc_hip_ctx **ctx);
/* print_layout of such a variant, without a context. */
int alpacc_hip_variant_layout(const char *grammar, int mode, int index32, int length_bits, char *buf, size_t cap);
void alpacc_hip_close(alpacc_hip_ctx *ctx);
const char *alpacc_hip_last_error(void);

/* Names of the grammars: those compiled in, then those loaded, NULL-terminated. */
const char *const *alpacc_hip_grammars(void);

/* Load a grammar's tables file at run time (the output of `python -m
 * alpacc_amd.gen hip GRAMMAR.alp -o FILE`): the counterpart of generating
 * one program per grammar with `alpacc c|cuda GRAMMAR.alp` (app/Main.hs:
 * 393-402, src/Alpacc/Generator/C/Generator.hs:161-207).  The grammar is
 * registered under its name (written to name, if not NULL) for the rest of
 * the process and opens like a compiled-in one.  Loading the same file
 * again is a no-op; another grammar of the same name is refused. */
int alpacc_hip_load_grammar(const char *path, char *name, size_t name_cap);

/* Size limits of one launch for a grammar, without a device: the input
 * bytes one lexer launch takes (alpacc_hip_run_batch and run_case split
 * larger batches into several launches; a larger lexer-only input goes
 * through the stream path, a larger combined one through the stream path
 * and then one parser launch over all its tokens) and the framed parser
 * positions (tokens + 2 per input) one parser launch is sure to take (an
 * input beyond them is refused with ALPACC_ERR_LIMIT). */
int alpacc_hip_grammar_limits(const char *grammar, uint64_t *launch_bytes, uint64_t *parser_positions);

/* Names of ids, as the reference's Futhark target names them in its
 * `terminal` and `production` types (src/Alpacc/Generator/Futhark/
 * Generator.hs:91-118, names from src/Alpacc/Generator/Analyzer.hs:104-110,
 * 158-165 and src/Alpacc/Encode.hs:334-341): a named terminal by its name,
 * a literal as literal_<id>, the dead terminal as empty_<id>; a production
 * by its [Name] in the grammar, else <lhs>_<id>, Literal_<id> or start_<id>.
 * name_count returns how many ids there are (or an error); name writes the
 * id's name into buf and returns its length. */
enum alpacc_hip_names {
  ALPACC_NAMES_TERMINAL_LEXER = 0, /* terminal ids of lexer-only mode */
  ALPACC_NAMES_TERMINAL = 1,       /* terminal ids of parser-only and combined modes */
  ALPACC_NAMES_PRODUCTION = 2
};
int alpacc_hip_name_count(const char *grammar, int kind);
int alpacc_hip_name(const char *grammar, int kind, uint64_t id, char *buf, size_t cap);

/* `--layout` (print_layout, src/Alpacc/Generator/C/Generator.hs:64-110):
 * key=value lines into buf; returns the length or a negative error. */
int alpacc_hip_layout(const alpacc_hip_ctx *ctx, int mode, char *buf, size_t cap);

/* INPUT_BYTES(n) of the mode (C/Generator.hs:64,87,117). */
uint64_t alpacc_hip_payload_bytes(const alpacc_hip_ctx *ctx, int mode, uint64_t n);

/* The same two answers by grammar name, without a device (a client can
 * size frames and print --layout before any GPU work). */
int alpacc_hip_grammar_layout(const char *grammar, int mode, char *buf, size_t cap);
uint64_t alpacc_hip_grammar_payload_bytes(const char *grammar, int mode, uint64_t n);

/* ---- host-memory entry points ------------------------------------------ */

/* lex_string (backends/c/lexer.c:39-97).  mode is LEXER or BOTH (terminal
 * numbering).  Arrays are malloc'd; release with alpacc_hip_free. */
int alpacc_hip_lex_string(alpacc_hip_ctx *ctx, int mode, const uint8_t *str, uint64_t n,
                          alpacc_hip_lexemes *out);

/* parse_test (backends/c/parser.c:49-98): n terminal ids (terminal_t) to the
 * leftmost derivation (production_t[*num_prods], malloc'd). */
int alpacc_hip_parse_test(alpacc_hip_ctx *ctx, const void *tokens, uint64_t n, void **prods,
                          uint64_t *num_prods);

/* Combined run: lex_string + parse_test + compute_parents + compact_tree
 * (backends/c/parser.c:103-153) -> tree; arrays malloc'd. */
int alpacc_hip_lex_parse(alpacc_hip_ctx *ctx, const uint8_t *str, uint64_t n, alpacc_hip_tree *out);

/* run_case (src/Alpacc/Generator/C/Generator.hs:64-152): one request payload
 * to one wire response record (u8 valid + fields), malloc'd. */
int alpacc_hip_run_case(alpacc_hip_ctx *ctx, int mode, uint64_t n, const uint8_t *payload,
                        uint8_t **record, uint64_t *record_len);

void alpacc_hip_free(void *p);

/* Timing with HIP events recorded on the launch stream (no reference
 * counterpart; the reference's CUDA CLI times with events in --benchmark,
 * backends/cuda/cli.cu:459-493).  Stages: "lex" = the lexer's launches,
 * "parse" = the parser's launches from the pair lookup to the tree; every
 * kernel is also timed by its own name (alpacc_hip_timed_kernels lists the
 * names seen, comma-separated).  set_timing resets the totals; while on,
 * each launch is bracketed by its own pair of events. */
int alpacc_hip_set_timing(alpacc_hip_ctx *ctx, int on);
int alpacc_hip_timing(const alpacc_hip_ctx *ctx, const char *stage, double *total_ms, uint64_t *runs);
const char *alpacc_hip_timed_kernels(const alpacc_hip_ctx *ctx);

/* Launch shape of the lexer.  The reference bakes BLOCK_SIZE and
 * ITEMS_PER_THREAD into the generated CUDA source from a per-architecture
 * table (backends/cuda/cli.cu:50-82, picked by the --sm-arch codegen
 * constant, src/Alpacc/Generator/Cuda/Generator.hs:77-92); here the shape is
 * a property of the context, measured per grammar: items_per_thread = input
 * bytes per lane (64 or 128), workgroups_per_cu = persistent workgroups of
 * the fold and emit kernels per CU (0 = as many as fit), lookback_window =
 * predecessor tiles examined per lookback probe (1..64); ms = the lexer's
 * time on the tuning sample.  autotune_lexer times every shape of {64, 128}
 * x {2, 3, 4, 0} x {4, 8, 16, 32, 64} on `reps` runs over a sample in device
 * memory (d_sample, n bytes), keeps the fastest in the context and returns
 * it; the stage timing totals are reset.  Every shape gives the same
 * results. */
typedef struct {
  int items_per_thread, workgroups_per_cu, lookback_window;
  double ms;
} alpacc_hip_lex_shape;

int alpacc_hip_set_lex_shape(alpacc_hip_ctx *ctx, const alpacc_hip_lex_shape *shape);
int alpacc_hip_get_lex_shape(const alpacc_hip_ctx *ctx, alpacc_hip_lex_shape *shape);
int alpacc_hip_autotune_lexer(alpacc_hip_ctx *ctx, int mode, const uint8_t *d_sample, uint64_t n, int reps,
                              alpacc_hip_lex_shape *best);

/* Tile shape of the parser (the reference's parser BLOCK_SIZE /
 * ITEMS_PER_THREAD, chosen per architecture when its CUDA program is
 * generated, backends/cuda/parser.cu:45-158, backends/cuda/cli.cu:50-82):
 * framed positions per tile, and positions per lane of the resolve pass,
 * so positions_per_tile / positions_per_lane lanes make its workgroup.
 * Shapes: (1024, 4) by default, (512, 4), (2048, 4), (1024, 2), (1024, 8);
 * every shape gives the same results.  autotune_parser lexes a device
 * sample (combined mode) and times the parser in every shape on `reps`
 * runs, keeps the fastest in the context and returns it (ms = the parser's
 * time on the sample); the stage timing totals are reset. */
typedef struct {
  int positions_per_tile, positions_per_lane;
  double ms;
} alpacc_hip_parse_shape;

int alpacc_hip_set_parse_shape(alpacc_hip_ctx *ctx, const alpacc_hip_parse_shape *shape);
int alpacc_hip_get_parse_shape(const alpacc_hip_ctx *ctx, alpacc_hip_parse_shape *shape);
int alpacc_hip_autotune_parser(alpacc_hip_ctx *ctx, const uint8_t *d_sample, uint64_t n, int reps,
                               alpacc_hip_parse_shape *best);

/* ---- device-memory entry points ---------------------------------------- */
/* Input already in device memory; results stay in device buffers owned by the
 * context, valid until the next call on it.  Only the count and verdict cross
 * to the host.  A context serves one call at a time. */

int alpacc_hip_lex_device(alpacc_hip_ctx *ctx, int mode, const uint8_t *d_str, uint64_t n,
                          alpacc_hip_lexemes *d_out, void *stream);

int alpacc_hip_parse_device(alpacc_hip_ctx *ctx, const void *d_tokens, uint64_t n, void **d_prods,
                            uint64_t *num_prods, void *stream);

int alpacc_hip_lex_parse_device(alpacc_hip_ctx *ctx, const uint8_t *d_str, uint64_t n,
                                alpacc_hip_tree *d_out, void *stream);

/* alpacc_hip_lex_device without allocating: no reference counterpart; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The lexer's scratch memory is a
 * static device array of ALPACC_NOALLOC_SCRATCH_BYTES on every device the
 * library runs on (held until the process ends); the outputs are the
 * caller's device arrays of `cap` entries (cap >= n: a token holds at least
 * one byte) -- terminal_t, int64 start, length_t -- and *count receives the
 * number of tokens.  Inputs above ALPACC_NOALLOC_MAX_INPUT bytes are refused
 * with ALPACC_ERR_LIMIT.  Results, return codes and the context's lexer
 * shape are those of alpacc_hip_lex_device.  Calls that overlap on one
 * device share its scratch array and race: serialise them.  The context
 * must have been opened (opening allocates its tables). */
#define ALPACC_NOALLOC_SCRATCH_BYTES (3u << 20)
#define ALPACC_NOALLOC_MAX_INPUT (64ull << 20)
int alpacc_hip_lex_device_noalloc(alpacc_hip_ctx *ctx, int mode, const uint8_t *d_str, uint64_t n, void *d_terminals,
                                  int64_t *d_starts, void *d_lengths, uint64_t cap, uint64_t *count, void *stream);

/* Lexing one input that arrives in chunks (the reference's chunked CUDA
 * lexer, LexerCtx and its update between chunks, backends/cuda/lexer.cu:
 * 224-431, and lexer.fut:75-196): begin, then push the chunks in order, the
 * last with is_last_chunk = 1 (it may be empty).  Each push lexes its chunk
 * in device memory from the state the earlier chunks left and returns, in
 * d_out (device arrays of the context, valid until its next call), the
 * tokens that close inside it; starts are offsets in the whole stream and a
 * token may start in an earlier chunk.  Concatenated, the pushes' tokens
 * are lex_string's for the whole input.  Returns 1 while (or, after the
 * last chunk, if) the input is accepted, 0 once it is rejected (a token
 * longer than length_t; after the last chunk also a final state that is not
 * accepting), < 0 on errors.  Chunks take fewer than 2^31 bytes each; the
 * stream has no limit. */
int alpacc_hip_lex_stream_begin(alpacc_hip_ctx *ctx, int mode);
int alpacc_hip_lex_stream_push(alpacc_hip_ctx *ctx, const uint8_t *d_chunk, uint64_t n, int is_last_chunk,
                               alpacc_hip_lexemes *d_out, void *stream);

/* ---- batches -------------------------------------------------------------
 * `count` inputs packed back to back: input i is bytes [offsets[i],
 * offsets[i + 1]) of the packed buffer (offsets: a HOST array of count + 1
 * entries; inputs need no alignment).  The whole batch runs as one launch
 * sequence: the lexer packs the inputs into shared tiles (128-byte chunks
 * per input), the parser gives every input tiles of its own.  The reference
 * answers a counted batch one
 * request at a time (generated run_case per frame, Generator/C/
 * Generator.hs:64-152, driven by backends/c/cli.c:167-189); these calls
 * replace that loop. */

/* Response records of every input, concatenated in input order; each is
 * byte-identical to alpacc_hip_run_case's record for that input alone.
 * Returns 0 or an error; *records is malloc'd (alpacc_hip_free). */
int alpacc_hip_run_batch(alpacc_hip_ctx *ctx, int mode, const uint8_t *packed, const uint64_t *offsets,
                         uint64_t count, uint8_t **records, uint64_t *records_len);

/* Lex + parse (combined mode) of a batch already in device memory.  Results
 * stay on the device; the per-input index arrays are host memory owned by
 * the context.  Valid until the next call on the context. */
typedef struct {
  uint64_t count;
  const uint8_t *valid;                  /* [count] 1 = accepted */
  const uint64_t *token_offsets;         /* [count + 1] first token of each input in tokens */
  const uint64_t *node_offsets;          /* [count] first node of each input in productions/parents */
  const uint64_t *node_counts;           /* [count] */
  const uint64_t *token_parent_offsets;  /* [count] first entry of each input in token_parents */
  alpacc_hip_lexemes tokens;             /* device; starts relative to their input */
  void *productions;                     /* device, production_t per node */
  int64_t *parents;                      /* device, relative to the input's first node */
  int64_t *token_parents;                /* device, relative to the input's first node */
} alpacc_hip_batch;

int alpacc_hip_lex_parse_batch_device(alpacc_hip_ctx *ctx, const uint8_t *d_packed, const uint64_t *offsets,
                                      uint64_t count, alpacc_hip_batch *out, void *stream);

/* Graphs for combined-mode batches (no reference counterpart: the
 * reference launches every kernel from the host, backends/cuda/cli.cu:
 * 236-274).  While on, alpacc_hip_run_batch (combined mode) and
 * alpacc_hip_lex_parse_batch_device run the lex -> parse -> tree sequence
 * as one HIP graph: it is sized for a capacity (inputs and bytes rounded up
 * to powers of two, never shrinking), planned on the device (no host
 * synchronisation between lexer and parser) and captured once; later
 * batches within the capacity that read the same input pointer on the same
 * stream replay it.  A batch beyond it, a moved workspace or another input
 * pointer captures again (the capture call first runs the sequence
 * eagerly).  Workspaces are sized for the capacity's worst case.  Per-kernel
 * timing does not cover graph runs.  graph_stats counts captures and
 * replays. */
int alpacc_hip_set_graphs(alpacc_hip_ctx *ctx, int on);
int alpacc_hip_graph_stats(const alpacc_hip_ctx *ctx, uint64_t *captures, uint64_t *replays);

/* Element width of the parser's per-lane bracket stacks in LDS (a variant of
 * the kernels, with the same results; no reference counterpart, as the lexer
 * shape above): 0 = automatic (1 byte when every bracket symbol of the
 * grammar is below 256, else 2), 1 = one byte (ALPACC_ERR_ARG when a symbol
 * is 256 or more), 2 = two bytes.  One-byte stacks take less LDS, so more
 * tiles are resolved at once per CU. */
int alpacc_hip_set_bracket_width(alpacc_hip_ctx *ctx, int bytes);

/* ---- serving ---------------------------------------------------------------
 * Continuous batching over resident grammars (no reference counterpart: the
 * reference's server answers one request at a time, backends/c/cli.c:
 * 167-189).  Every grammar gets `lanes` contexts (own stream, buffers and
 * worker thread); a free lane takes every request queued for its grammar in
 * one mode, up to `hbm_budget` bytes of estimated device memory admitted at
 * once (0 = 64 GiB) and within the device's free memory (hipMemGetInfo,
 * beyond what the lane's workspaces already hold, less a 1 GiB reserve and
 * what other lanes' admitted batches may still allocate), `max_batch`
 * requests (0 = no limit) and 8 MiB of payload (a larger request is a batch
 * of its own), and answers them with one batched launch sequence.  Records
 * are byte-identical to alpacc_hip_run_case's for each request alone. */
typedef struct alpacc_hip_server alpacc_hip_server;

int alpacc_hip_server_open(const char *const *grammars, int ngrammars, int device, int lanes, uint64_t hbm_budget,
                           uint64_t max_batch, alpacc_hip_server **out);
/* The same with every lane's context a width variant (alpacc_hip_open_variant). */
int alpacc_hip_server_open_variant(const char *const *grammars, int ngrammars, int device, int lanes,
                                   uint64_t hbm_budget, uint64_t max_batch, int index32, int length_bits,
                                   alpacc_hip_server **out);
void alpacc_hip_server_close(alpacc_hip_server *srv);

/* Asynchronous request (payload copied); *ticket names it for _wait. */
int alpacc_hip_server_submit(alpacc_hip_server *srv, int grammar, int mode, const uint8_t *payload, uint64_t n,
                             uint64_t *ticket);
/* Blocks until the request is answered; *record is malloc'd (alpacc_hip_free). */
int alpacc_hip_server_wait(alpacc_hip_server *srv, uint64_t ticket, uint8_t **record, uint64_t *record_len);

/* Open-loop replay for measurements: request i (grammar[i], mode[i], bytes
 * [offsets[i], offsets[i + 1]) of packed) is submitted arrival_ns[i] after
 * the start; latency_ns[i] = answer time - scheduled arrival.  The records
 * are concatenated in request order (*records malloc'd). */
int alpacc_hip_server_replay(alpacc_hip_server *srv, uint64_t count, const int32_t *grammar, const int32_t *mode,
                             const uint8_t *packed, const uint64_t *offsets, const uint64_t *arrival_ns,
                             uint64_t *latency_ns, uint8_t **records, uint64_t *records_len);

/* Combined-mode batches of every lane through its captured graph
 * (alpacc_hip_set_graphs). */
int alpacc_hip_server_set_graphs(alpacc_hip_server *srv, int on);

/* Batches run, requests answered and the largest batch so far. */
int alpacc_hip_server_stats(alpacc_hip_server *srv, uint64_t *batches, uint64_t *requests, uint64_t *max_batch);

#ifdef __cplusplus
}
#endif

#endif /* ALPACC_HIP_H */
