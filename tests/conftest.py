import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


GRAMMARS = ["arithmetic", "json", "sexp", "paper_grammar"]


@pytest.fixture(scope="session")
def analysed():
    from alpacc_amd.gen import tables
    out = {}
    for g in GRAMMARS:
        out[g] = tables.load(os.path.join(ROOT, "grammars", g + ".alp"))
    return out
