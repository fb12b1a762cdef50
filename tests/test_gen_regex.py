"""Regex front end + DFA construction against the reference's regex unit
tests (haskell-tests/RegularExpression.hs:40-124): every listed string must be
matched / not matched by the single-terminal DFA."""

import pytest

from alpacc_amd.gen import grammar as G
from alpacc_amd.gen import lexgen

CASES = [
    ("a*|b*", ["", "a", "aa", "aaa", "b", "bb", "bbb"], ["ab", "ba", "c", "aba", "bba", "xy"]),
    ("a*b*", ["", "ab", "aab", "b", "bb", "aaa", "abb", "aabbb"], ["aba", "bab", "xy", "abab", "bbaaa"]),
    ("a|b|c", ["a", "b", "c"], ["ab", "abc", "cba", "ac", "bb", "cc", "abcab", "xy", ""]),
    ("a+b+c+", ["abc", "aabbcc", "aaabbbccc", "aaaaaabbbbbbcccccc", "abbc"],
     ["ab", "acb", "bca", "aabccx", "aabcbccc", "abcabc", "xy", ""]),
    ("[a-z]+", ["abc", "aabbcc", "aaabbbccc", "aaaaaabbbbbbcccccc", "abbc", "ab", "acb", "bca", "aabccx",
                "aabcbccc", "abcabc", "xy"], ["", "4234", "324", "1", "<", "#", "\\"]),
    ("(a*b|c)d+(e|)", ["bd", "bde", "abd", "cd", "abdde", "cdde", "abde"],
     ["a", "b", "aa", "bb", "cc", "ade", "de", "xbdde", "abbddee"]),
    ("(xy*|z)ab+(c|)", ["xab", "xyab", "xyyab", "zab", "xabc", "xyabc", "xyyyabc"],
     ["x", "xy", "xz", "ab", "abb", "xyyy", "abc4", "zatbc"]),
    ("a*b+c*(d|)e+f", ["aabcdeef", "aabbcdeef", "abbcdef", "abbcccdeef", "abdeef"],
     ["aaaeef", "abcdf", "aabbcczdeef", "aabccddeef", "abccddef", "xy", ""]),
    ("(cat|dog|fish)+([0-9]|)", ["catfish", "catcatfish", "dogdogdog4", "fish3", "catdogfish0", "cat3"],
     ["catdoga", "0fishfish", "4cat4", " dogdogdogdog", "ddogfish3", "xy", ""]),
    ("()", [""], ["catdoga", "0fishfish", "4cat4", " dogdogdogdog", "ddogfish3", "xy"]),
    ("", [""], ["catdoga", "0fishfish", "4cat4", " dogdogdogdog", "ddogfish3", "xy"]),
]


def matches(dfa, s: bytes) -> bool:
    st = 1
    for b in s:
        if dfa.produce[st][b]:
            return False
        st = dfa.delta[st][b]
        if st == 0:
            return False
    return dfa.token[st] is not None


@pytest.mark.parametrize("regex,good,bad", CASES)
def test_regex_kat(regex, good, bad):
    r = G.parse_regex(regex)
    dfa = lexgen.build_token_dfa([(("T", "t"), r)], {("T", "t"): 0})
    for s in good:
        assert matches(dfa, s.encode()), (regex, s)
    for s in bad:
        assert not matches(dfa, s.encode()), (regex, s)


def test_regex_dialect():
    # '.' is a literal dot, \s is one space, whitespace inside /../ is skipped
    r = G.parse_regex(r"a . \s b")
    dfa = lexgen.build_token_dfa([(("T", "t"), r)], {("T", "t"): 0})
    assert matches(dfa, b"a. b")
    assert not matches(dfa, b"axb")
    assert not matches(dfa, b"a.b")
    assert G.produces_epsilon(G.parse_regex("a*"))
    assert not G.produces_epsilon(G.parse_regex("a+"))
    r = G.parse_regex(r"[\x41-\x43]x?")
    dfa = lexgen.build_token_dfa([(("T", "t"), r)], {("T", "t"): 0})
    assert matches(dfa, b"B") and matches(dfa, b"Cx") and not matches(dfa, b"D")
