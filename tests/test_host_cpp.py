"""The C++ host classes (code_mpm_amd/host, libmpm_physics.so) on the CPU:
the library and the bar driver are built and export the reference's class
members, and the bar world the C++ PhysicsEngine builds
(initializeWorld_Classic_Bar, through GridPoint_Factory and
MaterialPoint_Factory_Classic_CC) is the same, double for double, as the
bar scene the Python side builds (no device is touched)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "code_mpm_amd", "libmpm_physics.so")
DEMO = os.path.join(ROOT, "code_mpm_amd", "bin", "mpm_bar_demo")


def read_dump(path):
    raw = np.fromfile(path, dtype=np.float64)
    n, nr = raw[:2].view(np.int64)
    n, nr = int(n), int(nr)
    cols = raw[2:2 + 23 * n].reshape(23, n)
    reports = raw[2 + 23 * n:].reshape(nr, 9)
    return dict(x=cols[0:3], v=cols[3:6], sig=cols[6:12], F=cols[12:21], Ws=cols[21], V=cols[22]), reports


def test_host_library_exports_reference_members():
    assert os.path.exists(LIB) and os.path.exists(DEMO)
    out = subprocess.run(["nm", "-DC", LIB], capture_output=True, text=True, check=True).stdout
    for member in ("PhysicsEngine::runSimulation_Classic_DoublePass_MPLocks(double)",
                   "PhysicsEngine::runSimulation_Classic_DoublePass_MPLocks_Contact(double)",
                   "PhysicsEngine::runSimulation_Classic_SinglePass_MP(double)",
                   "PhysicsEngine::getMaterialPoints()", "PhysicsEngine::getGridPoints()",
                   "PhysicsEngine::reportConsole()", "GridPoint_Factory::createGrid",
                   "MaterialPoint_Factory_Classic_CC::createDomain_Cuboid",
                   "GridPoint_Mediator::findAdjacentGridPoints", "GridPoint_Mediator::calculateBases_Classic",
                   "ConstitutiveRelation::calculateIncrement_VonMisesHardening"):
        assert member in out, member


def test_bar_world_matches_python_scene(tmp_path):
    from code_mpm_amd import scenes
    path = str(tmp_path / "bar0.bin")
    subprocess.run([DEMO, "0", path], check=True, capture_output=True, timeout=60)
    got, reports = read_dump(path)
    sc = scenes.elastic_bar()
    assert got["x"].shape[1] == sc.n and reports.shape[0] == 0
    assert np.array_equal(got["x"], sc.fields["POSITION"])
    assert np.array_equal(got["v"], sc.fields["VELOCITY"])
    assert np.array_equal(got["V"], sc.fields["VOLUME"])
    F = got["F"]
    assert np.array_equal(F, sc.fields["DEFORMATION_GRADIENT"])
    assert not got["sig"].any() and not got["Ws"].any()


def test_driver_usage_without_arguments():
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
