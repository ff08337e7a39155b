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
CHECK = os.path.join(ROOT, "tests", "host", "_build", "mpm_host_check")


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
                   "PhysicsEngine::runSimulation_Phase(int, int, double)",
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


def mediator_positions(L, cells, rng):
    """Random interior points, points on node planes (both layers kept),
    points just outside either face (one layer), far outside and NaN (none)."""
    h = L / cells
    pts = [rng.uniform(0.0, 1.0, (200, 3)) * L]
    on = rng.uniform(0.0, 1.0, (60, 3)) * L
    axis = rng.integers(0, 3, 60)
    on[np.arange(60), axis] = rng.integers(0, cells[axis] + 1) * h[axis]
    pts.append(on)
    for ax in range(3):
        lo = rng.uniform(0.0, 1.0, (10, 3)) * L
        lo[:, ax] = -rng.uniform(0.0, 0.9, 10) * h[ax]
        hi = rng.uniform(0.0, 1.0, (10, 3)) * L
        hi[:, ax] = L[ax] + rng.uniform(0.0, 0.9, 10) * h[ax]
        pts += [lo, hi]
    far = rng.uniform(0.0, 1.0, (6, 3)) * L
    far[:3, 0] = 3.0 * L[0]
    far[3:, 1] = -2.0 * L[1]
    nan = rng.uniform(0.0, 1.0, (2, 3)) * L
    nan[:, 2] = np.nan
    return np.concatenate(pts + [far, nan])


@pytest.mark.parametrize("seed", range(3))
def test_mediator_matches_oracle(tmp_path, seed):
    """GridPoint_Mediator of the C++ host layer (findAdjacentGridPoints and
    calculateBases_Classic, GridPoint_Mediator.cpp:4-144) on the nodes the
    C++ GridPoint_Factory made: the same node list, in the same order, and
    the same shape values and gradients, double for double, as the oracle's
    mediator."""
    from oracle import oracle
    rng = np.random.default_rng(seed)
    L = rng.uniform(0.01, 2.0, 3)
    cells = rng.integers(1, 40, 3).astype(np.int32)
    x = mediator_positions(L, cells, rng)
    inp = np.concatenate([L, cells.astype(np.float64), [x.shape[0]], x.ravel()])
    src, dst = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    inp.tofile(src)
    subprocess.run([CHECK, "mediator", src, dst], check=True, capture_output=True, timeout=60)
    out = np.fromfile(dst, dtype=np.float64)
    h = L / cells
    pos, counts = 0, []
    for xp in x:
        cnt = int(out[pos])
        idx = out[pos + 1: pos + 9]
        pos += 9
        ref = oracle.find_adjacent(L, cells, xp)
        assert cnt == len(ref) and np.array_equal(idx[:cnt], ref) and np.all(idx[cnt:] == -1), xp
        nn = cells + 1
        for k, node in enumerate(ref):
            ix, iy, iz = node % nn[0], (node // nn[0]) % nn[1], node // (nn[0] * nn[1])
            xn = np.array([ix * L[0] / cells[0], iy * L[1] / cells[1], iz * L[2] / cells[2]])
            N, g = oracle.bases(h, xp, xn)
            got = out[pos + 4 * k: pos + 4 * k + 4]
            assert got[0] == N and np.array_equal(got[1:], g), (xp, node)
        pos += 4 * cnt
        counts.append(cnt)
    assert pos == out.size
    assert set(counts) >= {0, 4, 8}
