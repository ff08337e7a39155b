"""The BESO client on the GPU against the CPU checker: the smoothing filter
(mpm_eso_filter vs the reference's all-pairs loop, Canvas2D_CC.cpp:302-341)
and whole iterations of the loop of Main.cpp:204-333, with the designs run
as one batch of instances, bit for bit (particle summation order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(a, b) or bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("radius", [0.0012, 0.00085, 0.0003])
def test_eso_filter_matches_oracle(radius):
    from code_mpm_amd import beso, capi
    from oracle import oracle as orc
    cv = beso.l_problem(offset=2e-4)
    rng = np.random.default_rng(int(radius * 1e5))
    obj = rng.lognormal(0.0, 1.0, cv.n_voxels)
    obj[~cv.active] = -5.0
    g = capi.eso_filter(cv.active, cv.eso, obj, cv.nx, cv.ny, cv.d_Offset, radius)
    c = orc.eso_filter(cv.active, cv.eso, obj, cv.nx, cv.ny, cv.d_Offset, radius)
    assert _same(g, c)
    assert np.all(g[~cv.active] == -5.0)
    if radius < 2e-4:
        assert np.isnan(g[cv.active & ~cv.eso]).any()   # no ESO voxel within reach


def _oracle_simulate(scs):
    from oracle.oracle import OracleEngine
    out = []
    for sc in scs:
        o = OracleEngine(threads=1, order="particle")
        o.load_scene(sc)
        while o.run(0.5 * sc.console_interval) == 0:
            pass
        out.append((o.get_particle("ENERGY_STRAIN")[0].copy(), o.get_particle("ENERGY_STRAIN_MAX")[0].copy()))
        o.close()
    return out


def _designs():
    from code_mpm_amd import beso
    return [beso.Design(beso.l_problem(offset=0.002), radius=0.005),
            beso.Design(beso.l_problem(offset=0.002), radius=0.0035, ratio_rejection=0.1),
            beso.Design(beso.l_problem(offset=0.002), radius=0.006, ratio_redundant=0.1)]


@pytest.mark.parametrize("k", [1, 3])
def test_beso_iterations_match_oracle(k):
    from code_mpm_amd import beso
    from oracle import oracle as orc
    kw = dict(time_loading=2e-6)
    gpu = beso.BESO(_designs()[:k], order="particle", scene_kw=kw)
    cpu = beso.BESO(_designs()[:k], scene_kw=kw, simulate=_oracle_simulate, smooth=orc.eso_filter)
    for it in range(3):
        lg, lc = gpu.step(), cpu.step()
        assert lg == lc, f"iteration {it}: {lg} vs {lc}"
        for dg, dc in zip(gpu.designs, cpu.designs):
            assert np.array_equal(dg.canvas.opacity, dc.canvas.opacity), f"iteration {it}"
            assert _same(dg.canvas.objective, dc.canvas.objective), f"iteration {it}"
    assert any(d.removed > 0 for d in gpu.designs)
