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


# --- the wire-protocol program's framing (no GPU work happens before a valid
# --- frame, so these run on the CPU; cli.c:98-161 behaviour) ----------------

import struct

import pyoracle

CLI = os.path.join(ROOT, "alpacc_amd", "lib", "alpacc-hip")


def run_both(args, data: bytes, grammar="arithmetic", mode="both"):
    """Run our program and the oracle program on the same bytes."""
    flag = {"lexer": ["--lexer"], "parser": ["--parser"], "both": []}[mode]
    ours = subprocess.run([CLI, "-g", grammar] + flag + args, input=data, capture_output=True)
    ref = subprocess.run([pyoracle.program(grammar, mode)] + args, input=data, capture_output=True)
    return ours, ref


@pytest.mark.parametrize("grammar", ["arithmetic", "json", "sexp"])
@pytest.mark.parametrize("mode", ["lexer", "parser", "both"])
def test_cli_layout_matches_oracle(grammar, mode):
    if not os.path.exists(CLI):
        pytest.skip("cli not built")
    ours, ref = run_both(["--layout"], b"", grammar, mode)
    assert ours.returncode == 0 and ref.returncode == 0
    ref_kv = dict(l.split("=") for l in ref.stdout.decode().split())
    ours_kv = dict(l.split("=") for l in ours.stdout.decode().split())
    assert {k: ours_kv[k] for k in ref_kv} == ref_kv


FRAMING_CASES = {
    "empty stream": b"",
    "count only, no frame": struct.pack("<Q", 1),
    "zero requests": struct.pack("<Q", 0),
    "short frame length": struct.pack("<QQ", 1, 4),
    "truncated frame": struct.pack("<QQQ", 1, 20, 12) + b"1+2",
    "length mismatch": struct.pack("<QQQ", 1, 11, 5) + b"1+2",
    "truncated count": b"\x01\x00\x00",
}


@pytest.mark.parametrize("case", sorted(FRAMING_CASES))
def test_cli_framing_matches_oracle(case):
    if not os.path.exists(CLI):
        pytest.skip("cli not built")
    data = FRAMING_CASES[case]
    ours, ref = run_both([], data)
    assert ours.returncode == ref.returncode, (ours.stderr, ref.stderr)
    assert ours.stdout == ref.stdout
    assert ours.stderr.decode().strip() == ref.stderr.decode().strip()


def test_cli_server_empty_and_eof():
    if not os.path.exists(CLI):
        pytest.skip("cli not built")
    data = struct.pack("<Q", 0) * 3
    ours, ref = run_both(["--server"], data)
    assert ours.returncode == ref.returncode == 0
    assert ours.stdout == ref.stdout == data


def test_cli_invocation_by_link_name(tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("cli not built")
    link = tmp_path / "json_lexer"
    os.symlink(CLI, link)
    r = subprocess.run([str(link), "--layout"], capture_output=True)
    assert r.returncode == 0
    assert dict(l.split("=") for l in r.stdout.decode().split())["length_t"] == "2"
    assert subprocess.run([CLI, "-g", "nosuch", "--layout"], capture_output=True).returncode == 1


def test_wire_codec_roundtrip():
    from alpacc_amd import wire
    lay = wire.Layout("both", 1, 8, 2, 1)
    for payload in (b"", b"[1,2]", b"x" * 100):
        batch = wire.encode_batch([payload, payload], lay)
        cnt, fl, n = struct.unpack_from("<QQQ", batch, 0)
        assert cnt == 2 and fl == 8 + len(payload) and n == len(payload)
    rec = pyoracle.record("json", "both", b"[1,2]")
    r, off = wire.decode_record(rec, 0, lay)
    assert off == len(rec) and r.valid and r.terminals.tolist() == [4, 1, 2, 1, 5]
