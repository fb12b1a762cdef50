// Grammars loaded at run time: the binary tables file that
// alpacc_amd/gen/emit_hip.py writes for one grammar (the `alpacc hip`
// generator's output, in place of compiling the grammar into the library),
// and the registry that find_grammar searches before the compiled-in
// grammars.
//
// File layout (little-endian, every field padded to 8 bytes): the magic
// "ALPHIPT1", u32 field count, then the fields of GrammarDesc in declaration
// order -- the name as u32 length + bytes, a scalar as tag 'I' + i64, an
// array as tag 'A' + u32 entry size + u64 count + the entries.
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/alpacc_hip.h"
#include "grammar.hpp"

namespace alpacc {

int fail(int code, const char* fmt, ...);

namespace {

struct Reader {
  const uint8_t* p;
  size_t len, at = 0;
  bool ok = true;
  void align() { at = (at + 7) & ~size_t(7); }
  bool need(size_t n) {
    if (!ok || at > len || len - at < n) ok = false;
    return ok;
  }
  uint32_t u32() {
    uint32_t v = 0;
    if (need(4)) memcpy(&v, p + at, 4);
    at += 4;
    return v;
  }
  uint64_t u64() {
    uint64_t v = 0;
    if (need(8)) memcpy(&v, p + at, 8);
    at += 8;
    return v;
  }
};

// The fields of GrammarDesc in declaration order, as the file lists them.
enum Kind { K_INT, K_ARR };
struct Field {
  Kind kind;
  size_t off;   // offset in GrammarDesc
  int entry;    // array entry size
};

#define FI(f) Field{K_INT, offsetof(GrammarDesc, f), 0}
#define FA(f, e) Field{K_ARR, offsetof(GrammarDesc, f), e}
const Field FIELDS[] = {
    FI(q), FI(k), FI(nstate), FI(nendo), FI(nclass),
    FA(byte_class, 1), FA(step, 2), FA(compose, 2), FA(endo_init, 1), FA(delta_lex, 2), FA(delta_both, 2),
    FA(token_lex, 1), FA(token_both, 1),
    FI(ignore_lex), FI(ignore_both), FI(dead_lex), FI(dead_both),
    FI(terminal_bytes_lex), FI(terminal_bytes_both), FI(index_bytes), FI(length_bytes),
    FI(empty_t), FI(start_t), FI(end_t), FI(window_base),
    FI(nwindow), FA(window, 2), FI(npairs), FA(spans, 4), FA(profile, 1),
    FI(nstacks), FA(stacks, 2), FI(nprods_flat), FA(productions, 2),
    FI(nproductions), FI(production_bytes), FA(arity, 1), FA(prod_is_terminal, 1),
    FI(max_brackets), FI(max_prods), FI(max_slot_items), FA(pair_rec, 1), FA(prod_info, 1),
    FI(max_bpop), FI(max_bpush), FI(max_spop), FI(max_spush), FI(max_nt), FI(length_param), FI(has_parser),
    FA(terminal_names_lex, 1), FA(terminal_names_both, 1), FA(production_names, 1),
    FI(window_hash), FA(window_keys, 8), FA(pair_prod16, 2),
};
#undef FI
#undef FA

// Entries an array must hold for the kernels to stay inside it.
uint64_t need_entries(const GrammarDesc& g, size_t off) {
  auto n = [](int v) { return v > 0 ? (uint64_t)v : 1ull; };
  if (off == offsetof(GrammarDesc, byte_class)) return 256;
  if (off == offsetof(GrammarDesc, step)) return (uint64_t)g.nendo * g.nclass;
  if (off == offsetof(GrammarDesc, compose)) return (uint64_t)g.nendo * g.nendo;
  if (off == offsetof(GrammarDesc, endo_init)) return g.nendo;
  if (off == offsetof(GrammarDesc, delta_lex) || off == offsetof(GrammarDesc, delta_both)) return (uint64_t)g.nstate * 256;
  if (off == offsetof(GrammarDesc, token_lex) || off == offsetof(GrammarDesc, token_both)) return g.nstate;
  if (off == offsetof(GrammarDesc, window)) return n(g.nwindow);
  if (off == offsetof(GrammarDesc, spans)) return n(g.npairs * 4);
  if (off == offsetof(GrammarDesc, profile)) return n(g.npairs * 8);
  if (off == offsetof(GrammarDesc, stacks)) return n(g.nstacks);
  if (off == offsetof(GrammarDesc, productions)) return n(g.nprods_flat);
  if (off == offsetof(GrammarDesc, arity) || off == offsetof(GrammarDesc, prod_is_terminal) ||
      off == offsetof(GrammarDesc, prod_info))
    return n(g.nproductions);
  if (off == offsetof(GrammarDesc, pair_rec)) return (uint64_t)(g.npairs + 1) * 64;
  if (off == offsetof(GrammarDesc, window_keys)) return g.window_hash ? n(g.nwindow) : 1;
  if (off == offsetof(GrammarDesc, pair_prod16)) return g.production_bytes == 2 ? (uint64_t)(g.npairs + 1) * 8 : 1;
  return 0;
}

struct Loaded {
  std::vector<uint8_t> file;  // the tables file as read
  GrammarDesc desc{};
  std::string name;
  std::vector<std::vector<uint8_t>> storage;
};

std::mutex g_mu;
std::vector<std::unique_ptr<Loaded>> g_loaded;  // never freed: contexts keep pointers into them
std::vector<const char*> g_names;

}  // namespace

int parse_grammar_tables(const uint8_t* data, size_t len, GrammarDesc* out, std::vector<std::vector<uint8_t>>* storage,
                         std::string* name) {
  Reader r{data, len};
  if (len < 16 || memcmp(data, "ALPHIPT1", 8) != 0) return fail(ALPACC_ERR_ARG, "not an alpacc-hip tables file");
  r.at = 8;
  const uint32_t nfields = r.u32();
  r.align();
  const size_t nf = sizeof FIELDS / sizeof FIELDS[0];
  if (nfields != nf + 1) return fail(ALPACC_ERR_ARG, "tables file has %u fields, this library reads %zu", nfields, nf + 1);
  const uint32_t nl = r.u32();
  if (!r.need(nl) || nl == 0 || nl > 200) return fail(ALPACC_ERR_ARG, "tables file: bad grammar name");
  name->assign(reinterpret_cast<const char*>(data + r.at), nl);
  r.at += nl;
  r.align();
  GrammarDesc g{};
  std::vector<std::pair<size_t, uint64_t>> counts;  // arrays read: (field offset, entries)
  storage->clear();
  storage->reserve(nf);
  for (const Field& f : FIELDS) {
    if (!r.need(8)) break;
    const uint8_t tag = data[r.at];
    r.at += 8;
    if (f.kind == K_INT) {
      if (tag != 'I') return fail(ALPACC_ERR_ARG, "tables file: scalar expected at byte %zu", r.at - 8);
      const int64_t v = (int64_t)r.u64();
      if (v < -(1ll << 31) || v >= (1ll << 31)) return fail(ALPACC_ERR_ARG, "tables file: scalar out of range");
      *reinterpret_cast<int*>(reinterpret_cast<uint8_t*>(&g) + f.off) = (int)v;
    } else {
      if (tag != 'A') return fail(ALPACC_ERR_ARG, "tables file: array expected at byte %zu", r.at - 8);
      r.at -= 4;  // entry size follows the tag's first 4 bytes
      const uint32_t es = r.u32();
      const uint64_t cnt = r.u64();
      if ((int)es != f.entry || cnt > (1ull << 32) || !r.need(cnt * es))
        return fail(ALPACC_ERR_ARG, "tables file: bad array at byte %zu", r.at);
      const bool names = f.off == offsetof(GrammarDesc, terminal_names_lex) ||
                         f.off == offsetof(GrammarDesc, terminal_names_both) ||
                         f.off == offsetof(GrammarDesc, production_names);
      if (names && (cnt == 0 || data[r.at + cnt - 1] != 0))
        return fail(ALPACC_ERR_ARG, "tables file: names not NUL-terminated");
      storage->emplace_back(data + r.at, data + r.at + cnt * es);
      if (storage->back().empty()) storage->back().assign(8, 0);
      *reinterpret_cast<const void**>(reinterpret_cast<uint8_t*>(&g) + f.off) = storage->back().data();
      counts.emplace_back(f.off, cnt);
      r.at += cnt * es;
    }
    r.align();
  }
  if (!r.ok) return fail(ALPACC_ERR_ARG, "tables file truncated");
  // every array holds what the descriptor's sizes say the kernels read
  for (auto& [off, cnt] : counts)
    if (cnt < need_entries(g, off)) return fail(ALPACC_ERR_ARG, "tables file: an array is shorter than its sizes");
  if (g.nstate < 2 || g.nstate > 256 || g.nendo < 2 || g.nclass < 1 || g.q < 0 || g.k < 0 || g.q + g.k > 16)
    return fail(ALPACC_ERR_ARG, "tables file: sizes out of range");
  // the window table covers every code the parser forms: a dense index of
  // window_base^(q+k) entries, or a hashed table of 2^window_hash slots
  if (g.production_bytes < 1 || g.production_bytes > (g.has_parser ? 2 : 1))
    return fail(ALPACC_ERR_ARG, "tables file: %d-byte production ids", g.production_bytes);
  if (g.has_parser) {
    if (g.window_hash < 0 || g.window_hash > 30 || g.window_base < 1)
      return fail(ALPACC_ERR_ARG, "tables file: bad window table");
    if (g.window_hash) {
      if ((uint64_t)g.nwindow != (1ull << g.window_hash))
        return fail(ALPACC_ERR_ARG, "tables file: hashed window table of %d slots", g.nwindow);
    } else {
      uint64_t codes = 1;
      for (int i = 0; i < g.q + g.k && codes <= (uint64_t)g.nwindow; ++i) codes *= (uint64_t)g.window_base;
      if (codes > (uint64_t)g.nwindow) return fail(ALPACC_ERR_ARG, "tables file: window index shorter than its codes");
    }
  }
  *out = g;
  return 0;
}

const GrammarDesc* find_grammar(const char* name) {
  {
    std::lock_guard<std::mutex> lock(g_mu);
    for (auto& l : g_loaded)
      if (l->name == name) return &l->desc;
  }
  return find_builtin_grammar(name);
}

const char* const* grammar_names() {
  std::lock_guard<std::mutex> lock(g_mu);
  g_names.clear();
  for (const char* const* p = builtin_grammar_names(); *p; ++p) g_names.push_back(*p);
  for (auto& l : g_loaded) g_names.push_back(l->name.c_str());
  g_names.push_back(nullptr);
  return g_names.data();
}

}  // namespace alpacc

using namespace alpacc;

extern "C" int alpacc_hip_load_grammar(const char* path, char* name, size_t name_cap) {
  if (!path) return fail(ALPACC_ERR_ARG, "null argument");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(ALPACC_ERR_ARG, "cannot open tables file '%s'", path);
  std::vector<uint8_t> buf;
  uint8_t chunk[1 << 16];
  size_t got;
  while ((got = fread(chunk, 1, sizeof chunk, f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  fclose(f);
  auto l = std::make_unique<Loaded>();
  if (int r = parse_grammar_tables(buf.data(), buf.size(), &l->desc, &l->storage, &l->name)) return r;
  if (name && name_cap) {
    if (l->name.size() + 1 > name_cap) return fail(ALPACC_ERR_ARG, "name buffer too small");
    memcpy(name, l->name.c_str(), l->name.size() + 1);
  }
  std::lock_guard<std::mutex> lock(g_mu);
  for (auto& o : g_loaded)
    if (o->name == l->name) {
      // the same name again: the same tables are fine, others are refused
      return o->file == buf ? 0 : fail(ALPACC_ERR_ARG, "a different grammar named '%s' is already loaded", l->name.c_str());
    }
  if (find_builtin_grammar(l->name.c_str()))
    return fail(ALPACC_ERR_ARG, "grammar '%s' is compiled into the library", l->name.c_str());
  l->desc.name = l->name.c_str();
  l->file = std::move(buf);
  g_loaded.push_back(std::move(l));
  return 0;
}
