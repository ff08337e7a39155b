"""Writes tests/golden/oracle_runs.json: the CPU oracle's results on small
seeded workloads, as committed fixtures that pin the HIP path.

Each case names a seeded scene from code_mpm_amd.scenes (the inputs, which
the scene builders regenerate exactly from their seeds), the step loop, the
summation order, the strain formulation and the step count; the fixture
holds, per particle field, node field and monitor, a digest of the exact
doubles after the run (SHA-256 of the little-endian bytes with -0.0 written
as +0.0 and every NaN as one quiet NaN, which the parity tests also treat as
equal) and a few sampled values for reading, plus the health counters.  A
second part holds the smoothing filter of the BESO client on a seeded
canvas.

    python tests/golden/make_golden.py          # rewrites the fixture
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "oracle_runs.json")

PARTICLE_FIELDS = ["POSITION", "VELOCITY", "MASS", "VOLUME", "DEFORMATION_GRADIENT", "DEFORMATION_GRADIENT_ELASTIC",
                   "STRAIN", "STRESS", "STRAIN_PLASTIC", "BACKSTRESS", "ENERGY_STRAIN", "ENERGY_STRAIN_MAX",
                   "ENERGY_STRAIN_TI", "ENERGY_PLASTIC"]
NODE_FIELDS = ["MASS", "MOMENTUM", "VELOCITY", "FORCE", "FORCE_TEMP", "ACTIVE"]
HEALTH = ["steps", "nonfinite_particles", "outside_grid", "null_material", "plastic_points"]

# name: (scene builder, its arguments, loop, order, strain formulation, steps)
CASES = {
    "bar_doublepass_particle_100": ("elastic_bar", {}, "doublepass", "particle", "small", 100),
    "bar_doublepass_cell_100": ("elastic_bar", {}, "doublepass", "cell", "small", 100),
    "block0_doublepass_particle_12": ("random_block", {"seed": 0}, "doublepass", "particle", "small", 12),
    "block1_contact_particle_12": ("random_block", {"seed": 1}, "doublepass_contact", "particle", "small", 12),
    "block2_singlepass_particle_12": ("random_block", {"seed": 2}, "singlepass", "particle", "small", 12),
    "block3_doublepass_cell_12": ("random_block", {"seed": 3}, "doublepass", "cell", "small", 12),
    "block4_large_particle_10": ("random_block_fe", {"seed": 4}, "doublepass", "particle", "large", 10),
    "honeycomb64_doublepass_cell_60": ("honeycomb_crush", {"cells": [64, 64, 8], "steps": 300}, "doublepass", "cell",
                                       "small", 60),
    "halfring_doublepass_particle_30": ("half_ring_impact", {"h": 4.0e-4, "depth_cells": 4, "foam_cells": 8,
                                                             "impact_speed": 60.0, "steps": 90},
                                        "doublepass", "particle", "small", 30),
    "block5_doublepass_tile_12": ("random_block", {"seed": 5, "cells": [20, 18, 10], "n": 3000}, "doublepass",
                                  "tile", "small", 12),
    "block6_contact_tile_12": ("random_block", {"seed": 6, "cells": [18, 12, 9], "n": 2500}, "doublepass_contact",
                               "tile", "small", 12),
    "block7_singlepass_tile_12": ("random_block", {"seed": 7, "cells": [12, 20, 9], "n": 2500}, "singlepass",
                                  "tile", "small", 12),
    "honeycomb64_doublepass_tile_60": ("honeycomb_crush", {"cells": [64, 64, 8], "steps": 300}, "doublepass", "tile",
                                       "small", 60),
    "halfring_doublepass_tile_30": ("half_ring_impact", {"h": 4.0e-4, "depth_cells": 4, "foam_cells": 8,
                                                         "impact_speed": 60.0, "steps": 90},
                                    "doublepass", "tile", "small", 30),
    # the pair order (the default)
    "bar_doublepass_pair_100": ("elastic_bar", {}, "doublepass", "pair", "small", 100),
    "block8_doublepass_pair_12": ("random_block", {"seed": 8, "cells": [20, 18, 10], "n": 3000}, "doublepass",
                                  "pair", "small", 12),
    "block9_contact_pair_12": ("random_block", {"seed": 9, "cells": [18, 12, 9], "n": 2500}, "doublepass_contact",
                               "pair", "small", 12),
    "block10_singlepass_pair_12": ("random_block", {"seed": 10, "cells": [12, 20, 9], "n": 2500}, "singlepass",
                                   "pair", "small", 12),
    "block11_large_pair_10": ("random_block_fe", {"seed": 11}, "doublepass", "pair", "large", 10),
    "honeycomb64_doublepass_pair_60": ("honeycomb_crush", {"cells": [64, 64, 8], "steps": 300}, "doublepass", "pair",
                                       "small", 60),
    "halfring_doublepass_pair_30": ("half_ring_impact", {"h": 4.0e-4, "depth_cells": 4, "foam_cells": 8,
                                                         "impact_speed": 60.0, "steps": 90},
                                    "doublepass", "pair", "small", 30),
}


def build_scene(builder, kw):
    from code_mpm_amd import scenes
    if builder == "random_block_fe":
        # a random block with an elastic deformation gradient near the yield
        # strain, so that the large-strain trial states straddle the yield surface
        sc = scenes.random_block(**kw)
        rng = np.random.default_rng(100 + kw["seed"])
        E = sc.materials[0]["elastic_modulus"]
        Fe = np.eye(3)[:, :, None] + rng.normal(0.0, 60e6 / E, (3, 3, sc.n))
        sc.fields["DEFORMATION_GRADIENT_ELASTIC"] = Fe.transpose(1, 0, 2).reshape(9, sc.n)
        return sc
    return getattr(scenes, builder)(**{k: (tuple(v) if isinstance(v, list) else v) for k, v in kw.items()})


def canonical(a):
    a = np.asarray(a, np.float64)
    a = np.where(np.isnan(a), np.nan, a + 0.0)    # -0.0 -> +0.0, one NaN
    return np.ascontiguousarray(a, "<f8")


def digest(a):
    return hashlib.sha256(canonical(a).tobytes()).hexdigest()


def sample(a, idx):
    a = np.asarray(a, np.float64)
    return [[None if not np.isfinite(v) else float(v) for v in row] for row in a[:, idx]]


def summarize(engine, n, nn):
    rng = np.random.default_rng(12345)
    pidx = np.unique(np.concatenate([np.arange(min(4, n)), rng.integers(0, max(n, 1), 12)])) if n else []
    out = {"particles": {}, "nodes": {}}
    for f in PARTICLE_FIELDS:
        a = engine.get_particle(f)
        out["particles"][f] = {"sha256": digest(a)}
        if f in ("POSITION", "STRESS", "ENERGY_STRAIN"):
            out["particles"][f]["sample_index"] = [int(i) for i in pidx]
            out["particles"][f]["sample"] = sample(a, pidx)
    for f in NODE_FIELDS:
        out["nodes"][f] = {"sha256": digest(engine.get_node(f))}
    m = engine.monitors()
    out["monitors"] = {k: (None if isinstance(v, float) and not np.isfinite(v) else v) for k, v in m.items()}
    h = engine.health()
    out["health"] = {k: int(h[k]) for k in HEALTH}
    return out


def run_case(make_engine, spec):
    builder, kw, variant, order, strain, steps = spec
    sc = build_scene(builder, kw)
    e = make_engine(order, variant)
    e.load_scene(sc)
    e.set_strain_formulation(strain)
    e.step(sc.dt, steps)
    return summarize(e, sc.n, e.nnodes)


FILTER = {"size": [0.02, 0.02], "offset": 2.0e-4, "radius": 9.0e-4, "seed": 7}


def filter_case():
    from code_mpm_amd import beso
    cv = beso.l_problem(offset=FILTER["offset"], size=FILTER["size"])
    rng = np.random.default_rng(FILTER["seed"])
    obj = rng.lognormal(0.0, 1.0, cv.n_voxels)
    return cv, obj


def main():
    from oracle import oracle as orc
    from oracle.oracle import OracleEngine

    def make(order, variant):
        return OracleEngine(threads=1, order=order, variant=variant)

    res = {"about": "oracle results on seeded workloads; regenerate with tests/golden/make_golden.py",
           "cases": {}}
    for name, spec in CASES.items():
        res["cases"][name] = {"spec": {"builder": spec[0], "args": spec[1], "variant": spec[2], "order": spec[3],
                                       "strain": spec[4], "steps": spec[5]},
                              "result": run_case(make, spec)}
        print(name, res["cases"][name]["result"]["health"], flush=True)
    cv, obj = filter_case()
    out = orc.eso_filter(cv.active, cv.eso, obj, cv.nx, cv.ny, cv.d_Offset, FILTER["radius"])
    res["eso_filter"] = dict(FILTER, sha256=digest(out), nx=cv.nx, ny=cv.ny,
                             sample=[None if not np.isfinite(v) else float(v) for v in out[:: max(1, out.size // 16)]])
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
