"""The element routine's FP64 operation count (oracle/flopcount.cpp, a counting-double
build of the oracle) that the element kernel's roofline uses: the plastic branch costs
more than the elastic one, both lie in the range estimated for the
dense-as-written F-bar element (1e5 - 2e5 per element), and the count is reproducible."""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "..", "oracle", "flopcount")

ELEMENT_FLOPS_PLASTIC = 160073  # per Hex8 F-bar element, all 8 GPs plastic (bench.py uses this)
ELEMENT_FLOPS_ELASTIC = 155133


def _run():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.dirname(EXE), "flopcount"], check=True)
    out = subprocess.run([EXE], check=True, capture_output=True, text=True).stdout
    rows = dict(re.findall(r"^(\w+): .*flops per element (\d+)", out, re.M))
    return {k: int(v) for k, v in rows.items()}, out


def test_element_flop_count_is_stable_and_ordered():
    c1, out = _run()
    c2, _ = _run()
    assert c1 == c2
    assert "plastic Gauss points 8/8" in out and "plastic Gauss points 0/8" in out
    assert c1["elastic"] == ELEMENT_FLOPS_ELASTIC
    assert c1["plastic"] == ELEMENT_FLOPS_PLASTIC
    assert 1e5 < c1["elastic"] < c1["plastic"] < 2.5e5
