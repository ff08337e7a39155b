"""The C++ host classes on the GPU: the bar driver (PhysicsEngine::
initializeWorld_Classic_Bar, then runSimulation_Classic_DoublePass_MPLocks
over 250 steps' worth of time and getMaterialPoints) against the oracle
running the same world over the same request, bit for bit, in both
summation orders; the console reports fire at the reference's interval."""
import os
import subprocess

import numpy as np
import pytest

from tests.test_host_cpp import DEMO, read_dump

pytestmark = pytest.mark.gpu

STEPS = 250


@pytest.mark.parametrize("order", ["particle", "cell"])
def test_cpp_driver_matches_oracle(order, tmp_path):
    from code_mpm_amd import scenes
    from oracle.oracle import OracleEngine
    path = str(tmp_path / f"bar_{order}.bin")
    r = subprocess.run([DEMO, str(STEPS), path, "0" if order == "particle" else "1"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got, reports = read_dump(path)
    sc = scenes.elastic_bar()
    o = OracleEngine(threads=1, order=order)
    o.load_scene(sc)
    assert o.run(STEPS * sc.dt) == 0
    want = dict(x=o.get_particle("POSITION"), v=o.get_particle("VELOCITY"), sig=o.get_particle("STRESS"),
                F=o.get_particle("DEFORMATION_GRADIENT"), Ws=o.get_particle("ENERGY_STRAIN")[0],
                V=o.get_particle("VOLUME")[0])
    for k, w in want.items():
        assert np.array_equal(got[k], w), f"{order}: {k}"
    # reports at the first step whose start time is past 100 and 200 steps
    assert reports.shape[0] == 2
    assert 100 * sc.dt < reports[0, 0] < 102 * sc.dt and 200 * sc.dt < reports[1, 0] < 203 * sc.dt
    assert np.all(reports[:, 5] > 0.0)      # strain energy of the bending bar
