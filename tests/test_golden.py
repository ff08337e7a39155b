"""The oracle against the committed fixtures of tests/golden/oracle_runs.json
(written by tests/golden/make_golden.py): every case re-run on the CPU must
reproduce the digests of every particle and node field, the monitors and
the health counters, so a change to the oracle cannot pass unnoticed."""
import json
import os

import pytest

from tests.golden import make_golden as mg

FIXTURE = json.load(open(mg.OUT))


def _oracle(order, variant):
    from oracle.oracle import OracleEngine
    return OracleEngine(threads=1, order=order, variant=variant)


def _compare(got, want, label):
    for part in ("particles", "nodes"):
        for f, w in want[part].items():
            assert got[part][f]["sha256"] == w["sha256"], f"{label}: {part} {f}"
    assert got["health"] == want["health"], label
    for k, v in want["monitors"].items():
        assert got["monitors"][k] == v, f"{label}: monitor {k}"


@pytest.mark.parametrize("name", sorted(FIXTURE["cases"]))
def test_oracle_reproduces_fixture(name):
    spec = mg.CASES[name]
    _compare(mg.run_case(_oracle, spec), FIXTURE["cases"][name]["result"], name)


def test_oracle_filter_reproduces_fixture():
    from oracle import oracle as orc
    cv, obj = mg.filter_case()
    F = FIXTURE["eso_filter"]
    out = orc.eso_filter(cv.active, cv.eso, obj, cv.nx, cv.ny, cv.d_Offset, F["radius"])
    assert mg.digest(out) == F["sha256"]
