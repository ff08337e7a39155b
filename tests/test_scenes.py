"""Scene generators (CPU): geometry and boundary conditions of the C4
half-ring impact (PhysicsEngine_initializeWorld_Classic_HalfRing_Xiang_PlainStress.cpp)."""
import numpy as np

from code_mpm_amd import scenes
from code_mpm_amd.scenes import AXIS_X, AXIS_Y, AXIS_Z


def test_half_ring_impact_geometry():
    sc = scenes.half_ring_impact(h=4.0e-4, depth_cells=6, foam_cells=10)
    h = sc.h[0]
    x = sc.fields["POSITION"]
    n_ring = sc.info["ring_particles"]
    r_in, r_out = 0.5 * 0.0479, 0.5 * 0.0479 + 0.00148
    # ring particles: in the half annulus around (h, h + r_out), within a cell of the walls
    rr = np.hypot(x[0, :n_ring] - h, x[1, :n_ring] - (h + r_out))
    assert rr.min() > r_in - h and rr.max() < r_out + h
    assert x[0, :n_ring].min() >= h
    # it stands on the support: its lowest cells touch the fixed node row y = h
    assert x[1, :n_ring].min() < 2 * h
    # materials: steel ring, aluminium foam and platen
    assert (sc.material[:n_ring] == 0).all() and (sc.material[n_ring:] == 1).all()
    assert sc.materials[0]["yield_stress"] == 310.0e6 and sc.materials[1]["yield_stress"] == 190.0e6
    # foam above the ring, moving down; platen on top, controlled on all axes
    top = (x[1, :n_ring]).max()
    assert x[1, n_ring:].min() > top - h
    assert (sc.fields["VELOCITY"][1, n_ring:] == -20.0).all() and (sc.fields["VELOCITY"][:, :n_ring] == 0).all()
    assert (sc.dc_axes == (AXIS_X | AXIS_Y | AXIS_Z)).all() and len(sc.dc_index) == sc.info["platen_particles"]
    assert abs(sc.info["porosity"] - 0.5) < 0.05
    # nodes: every node fixed in z; x < 1.5 h fixed in x; y < 1.5 h fixed in y
    nodes = scenes.node_positions(sc.length, sc.cells)
    assert ((sc.fixed & AXIS_Z) != 0).all()
    assert (((sc.fixed & AXIS_X) != 0) == (nodes[0] < 1.5 * h)).all()
    assert (((sc.fixed & AXIS_Y) != 0) == (nodes[1] < 1.5 * h)).all()
    # masses from the two densities
    V0 = (h / 2) ** 3
    assert np.allclose(sc.fields["MASS"][:n_ring], 7800.0 * V0) and np.allclose(sc.fields["MASS"][n_ring:], 2760.0 * V0)


def test_half_ring_impact_default_size():
    """The default set-up is the ~48 M-particle workload (counted without building it)."""
    sc = scenes.half_ring_impact(h=2.0e-4, depth_cells=23, foam_cells=20)
    # the same body at 4 h: particles scale with (1/h)^2 in-plane and with depth, so the default
    # (h = 5e-5, 91 cells deep, 80 foam cells) has 16 x 91/23 times more
    est = sc.n * 16 * 91 / 23
    assert 40e6 < est < 56e6
