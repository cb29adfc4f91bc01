"""FP64 operation counts of the Hex8 F-bar element (oracle/flopcount.cpp, a counting-double
build): the oracle's routine as written (dense-as-written, 1e5 - 2.5e5 per element) and
the structure-exploiting form that the element kernel's roofline uses.  The structured
form must give the oracle's Fe and Ke (to rounding) on the elastic and the fully plastic
state; the plastic branch costs more than the elastic one; the counts are reproducible."""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "..", "oracle", "flopcount")

ELEMENT_FLOPS_PLASTIC = 160073  # per Hex8 F-bar element, all 8 GPs plastic (bench.py uses this)
ELEMENT_FLOPS_ELASTIC = 155133
STRUCTURED_PLASTIC = 83913  # bench.py's element roofline
STRUCTURED_ELASTIC = 80829


def _run():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.dirname(EXE), "flopcount"], check=True)
    out = subprocess.run([EXE], check=True, capture_output=True, text=True).stdout
    rows = dict(re.findall(r"^(\w+): .*?flops per element (\d+)", out, re.M))
    return {k: int(v) for k, v in rows.items()}, out


def test_element_flop_count_is_stable_and_ordered():
    c1, out = _run()
    c2, _ = _run()
    assert c1 == c2
    assert "plastic Gauss points 8/8" in out and "plastic Gauss points 0/8" in out
    assert c1["elastic"] == ELEMENT_FLOPS_ELASTIC
    assert c1["plastic"] == ELEMENT_FLOPS_PLASTIC
    assert 1e5 < c1["elastic"] < c1["plastic"] < 2.5e5


def test_structured_count_matches_the_oracle_element():
    c, out = _run()
    assert c["elastic_structured"] == STRUCTURED_ELASTIC
    assert c["plastic_structured"] == STRUCTURED_PLASTIC
    assert c["elastic_structured"] < c["plastic_structured"] < c["elastic"]
    diffs = re.findall(r"^(\w+_structured): .*max rel diff Fe (\S+) Ke (\S+)$", out, re.M)
    assert len(diffs) == 2
    for _, fe, ke in diffs:
        assert float(fe) < 1e-13 and float(ke) < 1e-13
    # the plastic state's tangent differs from the elastic one, so the Ke check above is
    # a check of the plastic branch too
    assert "plastic_structured: status 0, plastic Gauss points 8/8" in out
