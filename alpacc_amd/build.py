"""Build the HIP backend in-tree.

1. ``alpacc hip``: every grammar under grammars/ is analysed by
   alpacc_amd.gen and its device tables are emitted as
   alpacc_amd/csrc/generated/<grammar>.cpp, plus a registry source;
2. the HIP sources are compiled for gfx950 with hipcc and linked with the
   generated tables into alpacc_amd/lib/libalpacc_hip.so;
3. the wire-protocol program alpacc_amd/lib/alpacc-hip links that library.

Usage: python -m alpacc_amd.build [--jobs N]
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "generated")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
GRAMMARS = os.path.join(ROOT, "grammars")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
HIP_SOURCES = ["lexer.hip", "parse_seg.hip", "runtime.hip", "api_lex.hip", "api_parse.hip", "api_case.hip", "api_batch.hip", "serve.hip", "api_noalloc.hip", "tables_file.cpp"]

LIB_NAME = "libalpacc_hip.so"
CLI_NAME = "alpacc-hip"


def grammar_names():
    return sorted(os.path.splitext(f)[0] for f in os.listdir(GRAMMARS) if f.endswith(".alp"))


def _write_if_changed(path: str, text: str) -> None:
    if os.path.exists(path):
        with open(path) as f:
            if f.read() == text:
                return
    with open(path, "w") as f:
        f.write(text)


def generate() -> list:
    sys.path.insert(0, ROOT)
    from alpacc_amd.gen import emit_hip, tables

    os.makedirs(GEN, exist_ok=True)
    names = grammar_names()
    srcs = []
    for name in names:
        g = tables.load(os.path.join(GRAMMARS, name + ".alp"))
        path = os.path.join(GEN, f"{name}.cpp")
        _write_if_changed(path, emit_hip.emit(g))
        srcs.append(path)
    reg = ['// generated: registry of the compiled-in grammars', '#include <cstring>', '#include "../grammar.hpp"',
           "namespace alpacc {"]
    for name in names:
        reg.append(f"extern const GrammarDesc GRAMMAR_{name.upper()};")
    reg.append("static const GrammarDesc* const ALL[] = {" + ", ".join(f"&GRAMMAR_{n.upper()}" for n in names) + "};")
    reg.append("static const char* const NAMES[] = {" + ", ".join(f'"{n}"' for n in names) + ", nullptr};")
    reg += [
        "const GrammarDesc* find_builtin_grammar(const char* name) {",
        "  for (const GrammarDesc* g : ALL)",
        "    if (std::strcmp(g->name, name) == 0) return g;",
        "  return nullptr;",
        "}",
        "const char* const* builtin_grammar_names() { return NAMES; }",
        "}  // namespace alpacc",
        "",
    ]
    path = os.path.join(GEN, "registry.cpp")
    _write_if_changed(path, "\n".join(reg))
    srcs.append(path)
    return srcs


def _compile(src: str, obj: str, flags) -> str:
    cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def _stamp(path: str, flags) -> str:
    h = hashlib.sha256(" ".join(flags).encode())
    with open(path, "rb") as f:
        h.update(f.read())
    for dep in ("device.hpp", "lexer.hpp", "parser.hpp", "runtime.hpp", "grammar.hpp"):
        with open(os.path.join(CSRC, dep), "rb") as f:
            h.update(f.read())
    with open(os.path.join(ROOT, "include", "alpacc_hip.h"), "rb") as f:
        h.update(f.read())
    return h.hexdigest()[:16]


def build(jobs: int = 8, force: bool = True) -> str:
    """Compile everything; returns the library path."""
    gen_srcs = generate()
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    tasks = []
    for s in HIP_SOURCES:
        src = os.path.join(CSRC, s)
        tasks.append((src, HIP_FLAGS))
    for src in gen_srcs:
        tasks.append((src, ["-O1", "-fPIC", "-std=c++17", "-x", "c++"]))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = []
        for src, flags in tasks:
            obj = os.path.join(OBJ, os.path.basename(src) + "." + _stamp(src, flags) + ".o")
            objs.append(obj)
            if force or not os.path.exists(obj):
                futs.append(ex.submit(_compile, src, obj, flags))
        for f in futs:
            f.result()
    lib = os.path.join(LIB, LIB_NAME)
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(tmp, lib)
    cli = os.path.join(LIB, CLI_NAME)
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "cli.cpp"),
           "-o", cli + ".tmp", "-L", LIB, "-lalpacc_hip", "-Wl,-rpath,$ORIGIN"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(cli + ".tmp", cli)
    dg = os.path.join(LIB, "libdatagen.so")
    subprocess.run(["gcc", "-O2", "-std=c99", "-fPIC", "-shared", os.path.join(HERE, "datagen.c"), "-o", dg + ".tmp"],
                   check=True, capture_output=True, text=True)
    os.replace(dg + ".tmp", dg)
    return lib


def emit_test_tables(jobs: int = 8) -> list:
    """Tables files (python -m alpacc_amd.gen hip) of the parity tests'
    grammars under tests/golden/grammars/, written next to them as
    <name>.alphip; the tests load them into the library at run time."""
    sys.path.insert(0, ROOT)
    from alpacc_amd.gen import emit_hip, tables

    d = os.path.join(ROOT, "tests", "golden", "grammars")
    if not os.path.isdir(d):
        return []
    names = sorted(f for f in os.listdir(d) if f.endswith(".alp"))

    def one(f):
        out = os.path.join(d, f[:-4] + ".alphip")
        blob = emit_hip.emit_tables(tables.load(os.path.join(d, f)))
        if not os.path.exists(out) or open(out, "rb").read() != blob:
            with open(out, "wb") as fh:
                fh.write(blob)
        return out

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        return list(ex.map(one, names))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--incremental", action="store_true", help="reuse unchanged objects")
    a = ap.parse_args(argv)
    print(build(a.jobs, force=not a.incremental))
    print(f"{len(emit_test_tables(a.jobs))} test grammar tables")


if __name__ == "__main__":
    main()
