"""The C-ABI library loads and exports every entry point include/alpacc_hip.h
declares (no compute calls: this container has no GPU)."""

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "alpacc_hip.h")
LIB = os.path.join(ROOT, "alpacc_amd", "lib", "libalpacc_hip.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(alpacc_hip_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("alpacc_hip_open", "alpacc_hip_lex_string", "alpacc_hip_parse_test", "alpacc_hip_run_case",
                 "alpacc_hip_lex_device", "alpacc_hip_lex_parse_device", "alpacc_hip_layout"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_library_loads_with_ctypes():
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    from alpacc_amd import hip
    lib = hip.load_library()
    for n in declared():
        assert hasattr(lib, n)
    assert set(hip.grammars()) >= {"arithmetic", "json", "sexp"}


def test_cli_usage():
    cli = os.path.join(ROOT, "alpacc_amd", "lib", "alpacc-hip")
    if not os.path.exists(cli):
        pytest.skip("cli not built")
    r = subprocess.run([cli, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--server" in r.stderr
