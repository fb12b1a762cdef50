// alpacc-hip: the generated-program interface of the reference backends
// (backends/c/cli.c:1-189, docs/wire-protocols.md) over the HIP library.
//
//   alpacc-hip -g GRAMMAR [--lexer|--parser] [-i FILE] [-o FILE]
//              [--timeit] [--server] [--layout] [--device N]
//              [--index32] [--length N] [--names]
//   alpacc-hip --tables FILE [...]   a grammar generated at run time
//              (python -m alpacc_amd.gen hip GRAMMAR.alp -o FILE)
//
// --names prints the id -> name tables of the mode (the names the
// reference's Futhark target gives its terminal and production types) and
// exits.
// Invoked through a link named GRAMMAR or GRAMMAR_MODE (MODE = lexer, parser,
// both) it needs no -g / mode flag, so it can stand in for a generated
// program: batch mode reads one counted batch, --server loops over counted
// batches until EOF and flushes after every batch.  The frames of a batch go
// to the GPU together (alpacc_hip_run_batch).  Framing errors go to stderr
// with exit status 1, as in the reference.
#include <algorithm>
#include <chrono>
#include <exception>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "alpacc_hip.h"

namespace {

bool get_u64(FILE* f, uint64_t* v) { return fread(v, sizeof *v, 1, f) == 1; }
void put_u64(FILE* f, uint64_t v) { fwrite(&v, sizeof v, 1, f); }

struct Prog {
  alpacc_hip_ctx* ctx = nullptr;     // batch mode: opened at the first request
  alpacc_hip_server* srv = nullptr;  // --server: opened at the first request
  std::string grammar;
  int mode = ALPACC_MODE_BOTH;
  int device = 0;
  int index32 = 0, length_bits = 0;  // --index32, --length N
};

// Reads one frame's payload into `packed`.  0 = ok, -1 = framing error
// (reported on stderr, as backends/c/cli.c:98-143 does).
int read_frame(Prog& p, FILE* in, std::vector<uint8_t>& packed, std::vector<uint64_t>& off) {
  uint64_t flen;
  if (!get_u64(in, &flen)) {
    fprintf(stderr, "error: truncated input\n");
    return -1;
  }
  if (flen == UINT64_MAX || flen < 8) {
    fprintf(stderr, "error: bad frame length\n");
    return -1;
  }
  uint64_t n;
  if (!get_u64(in, &n)) {
    fprintf(stderr, "error: truncated frame\n");
    return -1;
  }
  // the payload is read in pieces of at most 64 MiB, so a bogus frame
  // length runs into the end of the input (a truncated frame) instead of
  // into one huge allocation
  const size_t at = packed.size();
  uint64_t left = flen - 8;
  while (left > 0) {
    const size_t piece = (size_t)std::min<uint64_t>(left, 64ull << 20), pos = packed.size();
    try {
      packed.resize(pos + piece);
    } catch (const std::exception&) {
      packed.resize(at);
      fprintf(stderr, "error: truncated frame\n");
      return -1;
    }
    const size_t got = fread(packed.data() + pos, 1, piece, in);
    if (got != piece) {
      packed.resize(at);
      fprintf(stderr, "error: truncated frame\n");
      return -1;
    }
    left -= piece;
  }
  if (flen != 8 + alpacc_hip_grammar_payload_bytes(p.grammar.c_str(), p.mode, n)) {
    fprintf(stderr, "error: frame length mismatch\n");
    return -1;
  }
  off.push_back(packed.size());
  return 0;
}

// Answers the frames read so far with one batch call.
int answer(Prog& p, const std::vector<uint8_t>& packed, const std::vector<uint64_t>& off, FILE* out) {
  if (off.size() < 2) return 0;
  if (!p.ctx && alpacc_hip_open_variant(p.grammar.c_str(), p.device, p.index32, p.length_bits, &p.ctx) < 0) {
    fprintf(stderr, "error: %s\n", alpacc_hip_last_error());
    return -1;
  }
  uint8_t* rec = nullptr;
  uint64_t len = 0;
  if (alpacc_hip_run_batch(p.ctx, p.mode, packed.data(), off.data(), off.size() - 1, &rec, &len) < 0) {
    fprintf(stderr, "error: %s\n", alpacc_hip_last_error());
    return -1;
  }
  fwrite(rec, 1, len, out);
  alpacc_hip_free(rec);
  return 0;
}

// --server: every frame goes to the serving runtime as soon as it is read,
// so requests of a batch are on the GPU while later ones still arrive; the
// records are written in request order.
int serve_batch(Prog& p, FILE* in, FILE* out, uint64_t count) {
  std::vector<uint8_t> packed;
  std::vector<uint64_t> off{0}, tickets;
  int rc = 0;
  for (uint64_t t = 0; t < count; ++t) {
    const size_t before = off.size();
    if (read_frame(p, in, packed, off) != 0) {
      rc = -1;
      break;
    }
    if (off.size() == before) continue;
    if (!p.srv) {
      const char* g = p.grammar.c_str();
      if (alpacc_hip_server_open_variant(&g, 1, p.device, 2, 0, 0, p.index32, p.length_bits, &p.srv) < 0) {
        fprintf(stderr, "error: %s\n", alpacc_hip_last_error());
        return -1;
      }
    }
    uint64_t ticket;
    const uint64_t a = off[off.size() - 2], b = off.back();
    if (alpacc_hip_server_submit(p.srv, 0, p.mode, packed.data() + a, b - a, &ticket) < 0) {
      fprintf(stderr, "error: %s\n", alpacc_hip_last_error());
      return -1;
    }
    tickets.push_back(ticket);
  }
  for (uint64_t t : tickets) {
    uint8_t* rec = nullptr;
    uint64_t len = 0;
    if (alpacc_hip_server_wait(p.srv, t, &rec, &len) < 0) {
      fprintf(stderr, "error: %s\n", alpacc_hip_last_error());
      return -1;
    }
    fwrite(rec, 1, len, out);
    alpacc_hip_free(rec);
  }
  return rc;
}

// 0 = batch done, 1 = clean EOF before a count (server), -1 = error.  The
// frames of a counted batch are read first and answered by one batch call;
// on a framing error the frames before it are answered, then the error is
// reported, so the output is the reference's frame-by-frame output.
int one_batch(Prog& p, FILE* in, FILE* out, bool server) {
  uint64_t count;
  if (!get_u64(in, &count)) {
    if (server && feof(in)) return 1;
    fprintf(stderr, "error: truncated input\n");
    return -1;
  }
  if (count == UINT64_MAX) {
    fprintf(stderr, "error: truncated input\n");
    return -1;
  }
  put_u64(out, count);
  if (server) return serve_batch(p, in, out, count);
  std::vector<uint8_t> packed;
  std::vector<uint64_t> off{0};
  for (uint64_t t = 0; t < count; ++t) {
    if (read_frame(p, in, packed, off) != 0) {
      // the message is already on stderr; answer what came before it
      fflush(stderr);
      if (answer(p, packed, off, out) != 0) return -1;
      return -1;
    }
  }
  return answer(p, packed, off, out);
}

int parse_mode(const std::string& s) {
  if (s == "lexer") return ALPACC_MODE_LEXER;
  if (s == "parser") return ALPACC_MODE_PARSER;
  if (s == "both") return ALPACC_MODE_BOTH;
  return -1;
}

void usage(const char* prog) {
  fprintf(stderr,
          "Usage: %s -g GRAMMAR [--lexer|--parser] [options]\n"
          "  -i FILE    input file  (default: stdin)\n"
          "  -o FILE    output file (default: stdout)\n"
          "  --timeit   print wall-clock elapsed time to stderr\n"
          "  --server   counted-batch loop until EOF (flush per batch)\n"
          "  --layout   print native type sizes and exit\n"
          "  --device N GPU ordinal (default 0)\n"
          "  --index32  32-bit index_t (and length_t when the grammar sets no length)\n"
          "  --length N length_t of N bits (8, 16, 32 or 64)\n"
          "  --tables FILE  load a grammar's tables file (alpacc_amd.gen hip)\n"
          "  --names    print the terminal and production names and exit\n"
          "  -h/--help  show this message\n",
          prog);
}

}  // namespace

int main(int argc, char** argv) {
  Prog p;
  std::string grammar;
  {
    std::string base = argv[0];
    size_t sl = base.rfind('/');
    if (sl != std::string::npos) base = base.substr(sl + 1);
    if (base != "alpacc-hip") {
      size_t us = base.rfind('_');
      if (us != std::string::npos && parse_mode(base.substr(us + 1)) >= 0) {
        p.mode = parse_mode(base.substr(us + 1));
        grammar = base.substr(0, us);
      } else {
        grammar = base;
      }
    }
  }
  const char *inp = nullptr, *outp = nullptr;
  bool timeit = false, server = false, layout = false, names = false;
  int device = 0;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-h" || a == "--help") {
      usage(argv[0]);
      return 0;
    } else if (a == "--timeit") {
      timeit = true;
    } else if (a == "--server") {
      server = true;
    } else if (a == "--layout") {
      layout = true;
    } else if (a == "--lexer" || a == "-l") {
      p.mode = ALPACC_MODE_LEXER;
    } else if (a == "--parser" || a == "-p") {
      p.mode = ALPACC_MODE_PARSER;
    } else if (a == "-g" && i + 1 < argc) {
      grammar = argv[++i];
    } else if (a == "-i" && i + 1 < argc) {
      inp = argv[++i];
    } else if (a == "-o" && i + 1 < argc) {
      outp = argv[++i];
    } else if (a == "--device" && i + 1 < argc) {
      device = atoi(argv[++i]);
    } else if (a == "--names") {
      names = true;
    } else if (a == "--tables" && i + 1 < argc) {
      char nm[256];
      if (alpacc_hip_load_grammar(argv[++i], nm, sizeof nm) < 0) {
        fprintf(stderr, "error: %s\n", alpacc_hip_last_error());
        return 1;
      }
      if (grammar.empty()) grammar = nm;
    } else if (a == "--index32") {
      p.index32 = 1;
    } else if (a == "--length" && i + 1 < argc) {
      p.length_bits = atoi(argv[++i]);
    } else {
      fprintf(stderr, "error: unknown argument '%s'\n", argv[i]);
      usage(argv[0]);
      return 1;
    }
  }
  if (grammar.empty()) {
    usage(argv[0]);
    return 1;
  }
  p.grammar = grammar;
  p.device = device;
  {
    char buf[512];
    if (alpacc_hip_variant_layout(grammar.c_str(), p.mode, p.index32, p.length_bits, buf, sizeof buf) < 0) {
      fprintf(stderr, "error: %s\n", alpacc_hip_last_error());
      return 1;
    }
    if (layout) {
      fputs(buf, stdout);
      return 0;
    }
  }
  if (names) {
    const int kinds[2] = {p.mode == ALPACC_MODE_LEXER ? ALPACC_NAMES_TERMINAL_LEXER : ALPACC_NAMES_TERMINAL,
                          ALPACC_NAMES_PRODUCTION};
    const char* label[2] = {"terminal", "production"};
    for (int k = 0; k < (p.mode == ALPACC_MODE_LEXER ? 1 : 2); ++k) {
      const int n = alpacc_hip_name_count(grammar.c_str(), kinds[k]);
      for (int id = 0; id < n; ++id) {
        char nm[256];
        if (alpacc_hip_name(grammar.c_str(), kinds[k], (uint64_t)id, nm, sizeof nm) < 0) {
          fprintf(stderr, "error: %s\n", alpacc_hip_last_error());
          return 1;
        }
        printf("%s %d %s\n", label[k], id, nm);
      }
    }
    return 0;
  }
  FILE* in = inp ? fopen(inp, "rb") : stdin;
  FILE* out = outp ? fopen(outp, "wb") : stdout;
  if (!in || !out) {
    perror(inp && !in ? inp : outp);
    return 1;
  }
  auto t0 = std::chrono::steady_clock::now();
  int rc = 0;
  if (server) {
    for (;;) {
      int r = one_batch(p, in, out, true);
      if (r == 1) break;
      if (r < 0) {
        rc = 1;
        break;
      }
      fflush(out);
    }
  } else {
    rc = one_batch(p, in, out, false) == 0 ? 0 : 1;
    fflush(out);
  }
  if (timeit) {
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    fprintf(stderr, "Time: %.2fms\n", ms);
  }
  if (in != stdin) fclose(in);
  if (out != stdout) fclose(out);
  if (p.ctx) alpacc_hip_close(p.ctx);
  if (p.srv) alpacc_hip_server_close(p.srv);
  return rc;
}
