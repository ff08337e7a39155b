"""The HIP path against the committed oracle fixtures of
tests/golden/oracle_runs.json: the same seeded workloads (regenerated by
the scene builders), every particle and node field bit for bit by digest,
the monitors and the health counters."""
import json

import pytest

from tests.golden import make_golden as mg

pytestmark = pytest.mark.gpu

FIXTURE = json.load(open(mg.OUT))


def _engine(order, variant):
    from code_mpm_amd import Engine
    return Engine(device=0, order=order, variant=variant)


@pytest.mark.parametrize("name", sorted(FIXTURE["cases"]))
def test_hip_path_matches_fixture(name):
    got = mg.run_case(_engine, mg.CASES[name])
    want = FIXTURE["cases"][name]["result"]
    for part in ("particles", "nodes"):
        for f, w in want[part].items():
            assert got[part][f]["sha256"] == w["sha256"], f"{name}: {part} {f}"
    assert got["health"] == want["health"], name
    for k, v in want["monitors"].items():
        assert got["monitors"][k] == v, f"{name}: monitor {k}"


def test_eso_filter_matches_fixture():
    from code_mpm_amd import capi
    cv, obj = mg.filter_case()
    F = FIXTURE["eso_filter"]
    out = capi.eso_filter(cv.active, cv.eso, obj, cv.nx, cv.ny, cv.d_Offset, F["radius"])
    assert mg.digest(out) == F["sha256"]
