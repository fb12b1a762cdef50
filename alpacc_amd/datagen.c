/*
 * Seeded synthetic inputs for the benchmark workloads (arithmetic expression
 * trees, concatenated JSON values, s-expression forests).
 *
 * Every generator fills exactly n bytes and is deterministic in (seed, n,
 * variant).  "literal" follows the workload descriptions as written (those
 * texts leave the reference grammars: JSON escapes and exponents, sexp atoms
 * with -+*, arithmetic "3-5" lexing as two numbers); "conforming" keeps every
 * byte inside the grammar so the whole lex + parse path does its work:
 *   - arithmetic: a space after a binary '-' that precedes a digit;
 *   - JSON: strings over [a-zA-Z0-9 ], no escapes or exponents, the values
 *     wrapped in one top-level array;
 *   - sexp: atoms over [a-zA-Z0-9], the forest wrapped in one list.
 * Padding up to n is whitespace, which every grammar ignores.
 *
 * PRNG: splitmix64 (Steele, Lea, Flood 2014).
 */
#include <stdint.h>
#include <string.h>

typedef struct {
  uint64_t s;
} rng_t;

static uint64_t next_u64(rng_t *r) {
  uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
/* uniform in [0, k) */
static uint32_t below(rng_t *r, uint32_t k) { return (uint32_t)(((next_u64(r) >> 32) * (uint64_t)k) >> 32); }
static double unit(rng_t *r) { return (double)(next_u64(r) >> 11) * (1.0 / 9007199254740992.0); }

typedef struct {
  uint8_t *p;
  uint64_t n, len;
  int full;
} buf_t;

static void put(buf_t *b, uint8_t c) {
  if (b->len < b->n)
    b->p[b->len++] = c;
  else
    b->full = 1;
}
static void puts_(buf_t *b, const char *s) {
  while (*s) put(b, (uint8_t)*s++);
}

static const char WS[4] = {' ', '\n', '\t', '\r'};

static void ws(buf_t *b, rng_t *r, uint32_t maxn) {
  uint32_t k = below(r, maxn + 1);
  for (uint32_t i = 0; i < k; i++) put(b, (uint8_t)WS[below(r, 4)]);
}

/* ---------------------------------------------------------------- arithmetic */

static int ar_expr(buf_t *b, rng_t *r, int depth, int conf);

static void ar_leaf_or_paren(buf_t *b, rng_t *r, int depth, int conf, int after_minus) {
  double u = unit(r);
  if (u < 0.1 && depth < 64) {
    put(b, '(');
    ws(b, r, 2);
    ar_expr(b, r, depth + 1, conf);
    ws(b, r, 2);
    put(b, ')');
    return;
  }
  if (u < 0.55) {
    int neg = unit(r) < 0.2;
    if (!neg && after_minus && conf) put(b, ' ');
    if (neg) put(b, '-');
    uint32_t d = 1 + below(r, 6);
    for (uint32_t i = 0; i < d; i++) put(b, (uint8_t)('0' + below(r, 10)));
    return;
  }
  uint32_t d = 1 + below(r, 8);
  for (uint32_t i = 0; i < d; i++) put(b, (uint8_t)('a' + below(r, 26)));
}

static void ar_term(buf_t *b, rng_t *r, int depth, int conf, int after_minus) {
  ar_leaf_or_paren(b, r, depth, conf, after_minus);
  while (!b->full && unit(r) >= 0.5) {
    ws(b, r, 2);
    put(b, below(r, 2) ? '*' : '/');
    ws(b, r, 2);
    ar_leaf_or_paren(b, r, depth, conf, 0);
  }
}

static int ar_expr(buf_t *b, rng_t *r, int depth, int conf) {
  ar_term(b, r, depth, conf, 0);
  while (!b->full && unit(r) >= 0.5) {
    ws(b, r, 2);
    int minus = below(r, 2);
    put(b, minus ? '-' : '+');
    uint64_t before = b->len;
    ws(b, r, 2);
    ar_term(b, r, depth, conf, minus && b->len == before);
  }
  return 0;
}

/* One expression of exactly n bytes: expressions joined by '+' while they
 * fit, then whitespace padding.  A candidate that does not fit is discarded
 * by rolling the cursor back. */
uint64_t datagen_arith(uint8_t *out, uint64_t n, uint64_t seed, int conforming) {
  rng_t r = {seed * 0x2545F4914F6CDD1Dull + 1};
  buf_t b = {out, n, 0, 0};
  int first = 1;
  for (;;) {
    uint64_t mark = b.len;
    if (!first) {
      put(&b, ' ');
      put(&b, '+');
      put(&b, ' ');
    }
    ar_expr(&b, &r, 0, conforming);
    if (b.full) {
      b.len = mark;
      b.full = 0;
      break;
    }
    first = 0;
    if (n - b.len < 8) break;
  }
  if (first && n > 0) { /* nothing fit: one number (or spaces) */
    b.len = 0;
    put(&b, '1');
  }
  while (b.len < n) out[b.len++] = ' ';
  return b.len;
}

/* ---------------------------------------------------------------------- JSON */

static void js_value(buf_t *b, rng_t *r, int depth, int conf);

static void js_string(buf_t *b, rng_t *r, int conf) {
  static const char ALNUM[] = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789 ";
  put(b, '"');
  uint32_t k = below(r, 33);
  for (uint32_t i = 0; i < k; i++) {
    if (conf) {
      put(b, (uint8_t)ALNUM[below(r, sizeof ALNUM - 1)]);
    } else if (unit(r) < 0.05) {
      static const char ESC[] = "\"\\/bfnrt";
      put(b, '\\');
      put(b, (uint8_t)ESC[below(r, sizeof ESC - 1)]);
    } else {
      uint8_t c;
      do c = (uint8_t)(32 + below(r, 95));
      while (c == '"' || c == '\\');
      put(b, c);
    }
  }
  put(b, '"');
}

static void js_number(buf_t *b, rng_t *r, int conf) {
  if (below(r, 2)) put(b, '-');
  uint32_t d = 1 + below(r, 10);
  for (uint32_t i = 0; i < d; i++) put(b, (uint8_t)('0' + (i == 0 && d > 1 ? 1 + below(r, 9) : below(r, 10))));
  if (unit(r) < 0.3) {
    put(b, '.');
    uint32_t f = 1 + below(r, 6);
    for (uint32_t i = 0; i < f; i++) put(b, (uint8_t)('0' + below(r, 10)));
  }
  if (!conf && unit(r) < 0.1) {
    put(b, 'e');
    if (below(r, 2)) put(b, '-');
    uint32_t e = 1 + below(r, 3);
    for (uint32_t i = 0; i < e; i++) put(b, (uint8_t)('0' + below(r, 10)));
  }
}

static void js_value(buf_t *b, rng_t *r, int depth, int conf) {
  double u = unit(r);
  if (b->full) return;
  if (u < 0.05) {
    puts_(b, "true");
  } else if (u < 0.10) {
    puts_(b, "false");
  } else if (u < 0.15) {
    puts_(b, "null");
  } else if (u < 0.25 && depth < 32) {
    put(b, '{');
    uint32_t k = below(r, 9);
    for (uint32_t i = 0; i < k && !b->full; i++) {
      if (i) put(b, ',');
      ws(b, r, 4);
      js_string(b, r, conf);
      ws(b, r, 4);
      put(b, ':');
      ws(b, r, 4);
      js_value(b, r, depth + 1, conf);
      ws(b, r, 4);
    }
    put(b, '}');
  } else if (u < 0.35 && depth < 32) {
    put(b, '[');
    uint32_t k = below(r, 9);
    for (uint32_t i = 0; i < k && !b->full; i++) {
      if (i) put(b, ',');
      ws(b, r, 4);
      js_value(b, r, depth + 1, conf);
      ws(b, r, 4);
    }
    put(b, ']');
  } else if (u < 0.65) {
    js_string(b, r, conf);
  } else {
    js_number(b, r, conf);
  }
}

/* wrap: one top-level array "[v,v,...]" (conforming parse input); else the
 * values are concatenated with whitespace between them. */
uint64_t datagen_json(uint8_t *out, uint64_t n, uint64_t seed, int conforming, int wrap) {
  rng_t r = {seed * 0x2545F4914F6CDD1Dull + 7};
  buf_t b = {out, n > 0 && wrap ? n - 1 : n, 0, 0};
  if (wrap) put(&b, '[');
  int first = 1;
  for (;;) {
    uint64_t mark = b.len;
    if (!first) {
      if (wrap) put(&b, ',');
      ws(&b, &r, 4);
    }
    js_value(&b, &r, wrap ? 1 : 0, conforming);
    if (b.full) {
      b.len = mark;
      b.full = 0;
      break;
    }
    first = 0;
  }
  if (wrap && n > 0) out[b.len++] = ']';
  while (b.len < n) out[b.len++] = ' ';
  return b.len;
}

/* ---------------------------------------------------------------------- sexp */

static void sx_atom(buf_t *b, rng_t *r, int conf) {
  static const char LIT[] = "abcdefghijklmnopqrstuvwxyz0123456789-+*/";
  static const char CONF[] = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789";
  uint32_t k = 1 + below(r, 12);
  for (uint32_t i = 0; i < k; i++)
    put(b, conf ? (uint8_t)CONF[below(r, sizeof CONF - 1)] : (uint8_t)LIT[below(r, sizeof LIT - 1)]);
}

/* A list nested `depth` levels: 0-16 children, one of which carries the
 * remaining depth, the others atoms. */
static void sx_list(buf_t *b, rng_t *r, uint32_t depth, int conf) {
  /* iterative: open every level, then close them, so depth 10000 is fine */
  uint32_t open = 0;
  for (uint32_t d = 0; d < depth && !b->full; d++) {
    put(b, '(');
    uint32_t k = below(r, 17);
    uint32_t carry = d + 1 < depth ? below(r, k + 1) : k + 1; /* position of the nested child */
    for (uint32_t i = 0; i < k && i < carry; i++) {
      if (i) put(b, ' ');
      sx_atom(b, r, conf);
    }
    if (d + 1 < depth) {
      if (carry) put(b, ' ');
      open++;
      continue;
    }
    put(b, ')');
  }
  for (uint32_t i = 0; i < open; i++) {
    uint32_t k = below(r, 4);
    for (uint32_t j = 0; j < k; j++) {
      put(b, ' ');
      sx_atom(b, r, conf);
    }
    put(b, ')');
  }
}

uint64_t datagen_sexp(uint8_t *out, uint64_t n, uint64_t seed, int conforming, int wrap) {
  rng_t r = {seed * 0x2545F4914F6CDD1Dull + 13};
  buf_t b = {out, n > 0 && wrap ? n - 1 : n, 0, 0};
  if (wrap) put(&b, '(');
  int first = 1;
  for (;;) {
    uint64_t mark = b.len;
    if (!first) put(&b, (uint8_t)WS[below(&r, 4)]);
    uint32_t depth;
    if (unit(&r) < 0.01) {
      depth = 10000;
    } else { /* geometric with mean 8 */
      depth = 1;
      while (unit(&r) >= 1.0 / 8.0) depth++;
    }
    if (unit(&r) < 0.3)
      sx_atom(&b, &r, conforming);
    else
      sx_list(&b, &r, depth, conforming);
    if (b.full) {
      b.len = mark;
      b.full = 0;
      break;
    }
    first = 0;
  }
  if (wrap && n > 0) out[b.len++] = ')';
  while (b.len < n) out[b.len++] = ' ';
  return b.len;
}
