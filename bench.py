"""Benchmark: particle-steps/s of the DoublePass MPM timestep on one MI355X.

Workload (the config the headline metric is quoted on): J2 honeycomb crush,
512 x 512 x 32 cells, 8.0 M particles, von Mises hardening with the exact
FP64 return map, displacement-controlled top platen at 10 m/s, fixed-y
support.  Synthetic seeded input (seed 1), resident in HBM before timing.

Prints one JSON line.  ``roofline`` is for the dominant kernel: algorithmic
bytes per launch over its average launch time from HIP events on the
engine's stream.  ``cpu_baseline`` is the CPU oracle (the reference's
OpenMP lock-scatter structure) timed on this host on a slice of the same
scene.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 8M J2 particles; surrogate vs exact speedup at stated err"
PEAK_HBM_GBS = 8000.0

# Algorithmic bytes per kernel (FP64 words counted once per pass that needs
# them; intermediate records and the binning are implementation traffic):
#   p2g_partials: reads x3 v3 m V sigma6 (14 words, +3 with external force)
#   p2g_nodes: per active node writes m p3 f3 (7 words, the P2G flush) and
#     reads fixed/friction and writes p3 f3 f/m3 reaction (13, grid update)
#   g2p_velocity: reads x3 v3, writes v3; per active node reads f/m3 m
#   remap_partials: reads x3 v3 m (7 words)
#   remap_nodes: per active node reads m p3, writes v3 p/m3 (10 words)
#   g2p_kinematics: reads x3 F9 V0, writes x3 F9 V (26 words) + material id
#     and control flags (5 B); per active node reads p/m3 v3 m (7 words)
#   g2p_constitutive: reads sigma6 eps6 R Ws Wti Wmax and writes sigma6
#     eps6 Ws Wti (30 words); a point that yields also reads and writes
#     epsp6 Wp and writes R (15 words; elsewhere they are unchanged)
PER_PARTICLE = {"p2g_partials": 14 * 8, "p2g_nodes": 0, "g2p_velocity": 9 * 8, "remap_partials": 7 * 8,
                "remap_nodes": 0, "g2p_kinematics": 26 * 8 + 5, "g2p_constitutive": 30 * 8, "bin": 0}
PER_PLASTIC = {"g2p_constitutive": 15 * 8, "g2p_kinematics": 15 * 8}
PER_NODE = {"p2g_partials": 0, "p2g_nodes": 20 * 8, "g2p_velocity": 4 * 8, "remap_partials": 0,
            "remap_nodes": 10 * 8, "g2p_kinematics": 7 * 8, "g2p_constitutive": 0, "bin": 0}
# With the exact return map, G2P pass 2 is one kernel
# (k_g2p_fused), booked in the kinematics slot: its algorithmic bytes are
# the kinematics' and the state update's together (the strain increment no
# longer crosses memory), and the constitutive slot launches nothing.
FUSED_G2P = {"g2p_kinematics": (26 * 8 + 5 + 30 * 8, 7 * 8)}
# kernel of each slot, per summation order (names as rocprofv3 reports them)
KERNELS = {
    "cell": {"bin": "k_scan_reduce+k_scan_blocks+k_scan_apply+k_scatter+k_cell_sort",
             "p2g_partials": "k_cell_partials<false, false, false, false>",
             "p2g_nodes": "k_node_cells<false, false, false>", "g2p_velocity": "k_g2p_velocity<false>",
             "remap_partials": "k_cell_partials<true, false, false, false>",
             "remap_nodes": "k_node_cells<true, false, false>", "g2p_kinematics": "k_g2p_fused<0, false>",
             "g2p_constitutive": None},
    "pair": {"bin": "k_scan_reduce+k_scan_blocks+k_scan_apply+k_scatter+k_cell_sort",
             "p2g_partials": "k_cell_partials<false, false, false, true>",
             "p2g_nodes": "k_node_cells<false, false, true>", "g2p_velocity": "k_g2p_velocity<false>",
             "remap_partials": "k_cell_partials<true, false, false, true>",
             "remap_nodes": "k_node_cells<true, false, true>", "g2p_kinematics": "k_g2p_fused<0, false>",
             "g2p_constitutive": None},
    "tile": {"bin": "k_scan_reduce+k_scan_blocks+k_scan_apply+k_scatter+k_cell_sort",
             "p2g_partials": "k_tile_p2g<false, false, false>", "p2g_nodes": "k_tile_bnd<false, false>",
             "g2p_velocity": "k_g2p_velocity<false>", "remap_partials": "k_tile_p2g<true, false, false>",
             "remap_nodes": "k_tile_bnd<true, false>", "g2p_kinematics": "k_g2p_fused<0, false>",
             "g2p_constitutive": None},
    "particle": {"bin": "k_scan_reduce+k_scan_blocks+k_scan_apply+k_scatter+k_cell_sort",
                 "p2g_partials": "k_list_terms<false, false, false>",
                 "p2g_nodes": "k_p2g_list<false>", "g2p_velocity": "k_g2p_velocity<false>",
                 "remap_partials": "k_list_terms<true, false, false>", "remap_nodes": "k_remap_list",
                 "g2p_kinematics": "k_g2p_fused<0, false>", "g2p_constitutive": None},
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_honeycomb8m.txt")


def read_traffic(path=TRAFFIC_FILE):
    """HBM bytes per launch per kernel, from a rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by a copy of known size; by default the
    committed one)."""
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        if line.startswith("#") or "hbm_bytes_per_launch[" not in line:
            continue
        tail = line.split("hbm_bytes_per_launch[", 1)[1]
        name, val = tail.split("]:", 1)
        out[name.strip()] = float(val)
    return out


def cpu_baseline(threads, steps=300, warmup=3):
    """The oracle (the reference's structure: AoS particles, per-node locks,
    per-step adjacency rebuild) on the x < 64 h slice of the workload; 300
    steps are about 10 s of work at 16 threads.  It runs in a child process
    that touches no GPU, with its OpenMP team pinned one thread per core
    (OMP_PROC_BIND=close, OMP_PLACES=cores); returns (particle-steps/s,
    particles, steps, seconds, the child's view of the host)."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child",
                        f"{threads},{steps},{warmup}"], env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed ({r.returncode}): {r.stderr[-2000:]}")
    d = json.loads(r.stdout.strip().splitlines()[-1])
    return d["value"], d["particles"], d["steps"], d["seconds"], d["host"]


def _cpu_baseline_child(spec):
    threads, steps, warmup = (int(v) for v in spec.split(","))
    # read before the OpenMP team starts: binding pins this thread to one core
    host = {"logical_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}
    from code_mpm_amd import scenes
    from oracle.oracle import OracleEngine
    sample = scenes.crop_x(scenes.honeycomb_crush(), 64)
    o = OracleEngine(threads=threads)
    o.load_scene(sample)
    o.step(sample.dt, warmup)
    t0 = time.perf_counter()
    o.step(sample.dt, steps)
    el = time.perf_counter() - t0
    o.close()
    print(json.dumps({"value": sample.n * steps / el, "particles": sample.n, "steps": steps, "seconds": el,
                      "host": host}), flush=True)


def batch_bench(n_inst, steps):
    """Graded-foam batch: n_inst independent instances (seeds 1000 + i) in one
    engine, steps replayed from one captured HIP graph."""
    from code_mpm_amd import Engine, scenes
    t = time.perf_counter()
    scs = [scenes.graded_foam(1000 + i) for i in range(n_inst)]
    eng = Engine(device=0, order="pair")
    eng.load_batch(scs)
    ntot = sum(sc.n for sc in scs)
    print(f"# batch: {n_inst} instances, {ntot} particles, setup {time.perf_counter() - t:.1f} s",
          file=sys.stderr, flush=True)
    eng.step_graph(steps)          # capture and first replay
    eng.synchronize()
    t0 = time.perf_counter()
    eng.step_graph(steps)
    eng.synchronize()
    el = time.perf_counter() - t0
    eng.close()
    return {"workload": f"graded J2 foam batch, {n_inst} instances (seeds 1000..{999 + n_inst}), "
                        "perfectly plastic, porosity U[0.3, 0.7], 1e4-2e5 particles each",
            "instances": n_inst, "particles": ntot, "steps": steps, "graph": True,
            "value": round(ntot * steps / el, 1), "unit": "particle-steps/s",
            "ms_per_step": round(1e3 * el / steps, 4)}


def surrogate_bench(steps, seeds, capture_steps, rows_total, train_steps, stress_steps=10, tol=1e-4):
    """C3: the learned return map on the J2 honeycomb crush.

    Training data: return-map samples the exact FP64 map itself produces,
    captured (mpm_set_sample_capture) from full-size runs of the same
    generator with other seeds, up to rows_total / len(seeds) rows each.
    Evaluation: the seed-1 crush, `steps` steps with the exact map and
    `steps` with the learned one (residual tolerance `tol` sigma_y, and once
    more at tol / 100), all from the same state; step time, the share of
    network answers accepted and the stress error at the end.  A second, stress-test comparison puts
    every particle's stress on or near the yield surface, so that every
    point takes the plastic branch, and times `stress_steps` steps."""
    import numpy as np
    from code_mpm_amd import Engine, scenes
    from code_mpm_amd import surrogate as su
    log = lambda m: print(f"# c3 {m}", file=sys.stderr, flush=True)
    t = time.perf_counter()
    rows, info = su.capture_from_runs(seeds, capture_steps, rows_total // len(seeds), log=log)
    t_capture = time.perf_counter() - t
    scene = scenes.honeycomb_crush(seed=1)
    mat = scene.materials[0]
    rows = rows[rows[:, 13] == 0]
    t = time.perf_counter()
    data = su.make_training_set(mat, rows[:, :13])
    nrows = int(rows.shape[0])
    del rows
    s = su.Surrogate(seed=11)
    s.fit_normalization(data)
    xs, ts = s.standardise(data)
    ntrain = int(xs.shape[0])
    del data
    t_prep = time.perf_counter() - t
    t = time.perf_counter()
    hist = s.train(xs, ts, steps=train_steps, batch=4096, lr=2e-3, seed=12)
    t_train = time.perf_counter() - t
    del xs, ts
    log(f"{nrows} captured rows, {ntrain} yielding, capture {t_capture:.1f} s, prep {t_prep:.1f} s, "
        f"train {t_train:.1f} s")

    def compare(sc, nsteps, control=False, tight=None):
        out, stress = {}, {}
        for mode in ("exact", "surrogate") + (("control",) if control else ()) + (("tight",) if tight else ()):
            # control: the exact map with the tile summation order, a change of
            # round-off size only, to show how far the crush's own dynamics
            # carry a perturbation over the same steps; tight: the same
            # network with a tighter residual tolerance
            eng = Engine(device=0, order="tile" if mode == "control" else "pair")
            eng.load_scene(sc)
            if mode in ("surrogate", "tight"):
                eng.set_surrogate(s, material=0, residual_tolerance=tol if mode == "surrogate" else tight)
            # both runs take the same two untimed steps (first-step allocations
            # stay out of the window) and then the timed ones
            eng.step(sc.dt, 2)
            eng.synchronize()
            t0 = time.perf_counter()
            eng.step(sc.dt, nsteps)
            eng.synchronize()
            el = time.perf_counter() - t0
            out[mode] = {"ms_per_step": round(1e3 * el / nsteps, 4), "health": eng.health()}
            stress[mode] = eng.get_particle("STRESS")
            eng.close()
            log(f"{sc.name} {mode}: {out[mode]['ms_per_step']:.3f} ms/step")
        a, b = stress["exact"], stress["surrogate"]
        rel = np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(a, axis=0), mat["yield_stress"])
        h = out["surrogate"]["health"]
        calls, falls = h["surrogate_calls"], h["surrogate_fallbacks"]
        extra = {}
        if control:
            c = stress["control"]
            rc = np.linalg.norm(a - c, axis=0) / np.maximum(np.linalg.norm(a, axis=0), mat["yield_stress"])
            extra["order_control"] = {"what": "exact map, tile summation order, against the exact pair-order run",
                                      "stress_rel_error_max": float(rc.max()),
                                      "stress_rel_error_mean": float(rc.mean()),
                                      "stress_rel_error_l2": float(np.linalg.norm(a - c) / np.linalg.norm(a))}
        if tight:
            t = stress["tight"]
            rt = np.linalg.norm(a - t, axis=0) / np.maximum(np.linalg.norm(a, axis=0), mat["yield_stress"])
            ht = out["tight"]["health"]
            extra["tighter_tolerance"] = {
                "residual_tolerance_sigma_y": tight, "surrogate_ms_per_step": out["tight"]["ms_per_step"],
                "speedup": round(out["exact"]["ms_per_step"] / out["tight"]["ms_per_step"], 3),
                "network_calls": ht["surrogate_calls"],
                "accepted_fraction": round((ht["surrogate_calls"] - ht["surrogate_fallbacks"]) /
                                           max(ht["surrogate_calls"], 1), 4),
                "stress_rel_error_max": float(rt.max()), "stress_rel_error_mean": float(rt.mean()),
                "stress_rel_error_l2": float(np.linalg.norm(a - t) / np.linalg.norm(a))}
        return {**extra, "steps": nsteps, "exact_ms_per_step": out["exact"]["ms_per_step"],
                "surrogate_ms_per_step": out["surrogate"]["ms_per_step"],
                "speedup": round(out["exact"]["ms_per_step"] / out["surrogate"]["ms_per_step"], 3),
                "plastic_points_exact": out["exact"]["health"]["plastic_points"],
                "network_calls": calls, "accepted_fraction": round((calls - falls) / max(calls, 1), 4),
                "stress_rel_error_max": float(rel.max()), "stress_rel_error_mean": float(rel.mean()),
                "stress_rel_error_l2": float(np.linalg.norm(a - b) / np.linalg.norm(a))}

    crush = compare(scene, steps, control=True, tight=tol / 100.0)
    rows = su.synthetic_states(mat, scene.n, seed=7)
    scene.fields["STRESS"] = np.ascontiguousarray(rows[:, :6].T)
    scene.fields["BACKSTRESS"] = np.ascontiguousarray(rows[:, 12])
    del rows
    stress_test = compare(scene, stress_steps)
    return {"workload": f"J2 honeycomb crush, 512x512x32 cells, {scene.n} particles, seed 1: exact FP64 return map "
                        f"against the learned one (residual tolerance {tol:g} sigma_y, stress error relative to "
                        "max(|sigma|, sigma_y))",
            "training": {"source": "mpm_set_sample_capture on full-size crushes of the same generator",
                         "seeds": [i[0] for i in info], "steps_run": [i[1] for i in info],
                         "rows_captured": nrows, "rows_yielding": ntrain, "train_steps": train_steps,
                         "batch": 4096, "capture_s": round(t_capture, 1), "prep_s": round(t_prep, 1),
                         "train_s": round(t_train, 2),
                         "loss_first_last": [float(hist[:10].mean()), float(hist[-10:].mean())]},
            "residual_tolerance_sigma_y": tol,
            "crush": crush,
            "all_yielding": dict(stress_test, workload="the same particles with every stress put on or near the "
                                                       "yield surface (every point takes the plastic branch)")}


def _train_surrogate(mat, train_steps):
    from code_mpm_amd import surrogate as su
    s = su.Surrogate(seed=11)
    data = su.make_training_set(mat, su.synthetic_states(mat, 200000, seed=3))
    s.fit_normalization(data)
    xs, ts = s.standardise(data)
    hist = s.train(xs, ts, steps=train_steps, batch=4096, lr=2e-3, seed=12)
    return s, hist


def _run_curve(scene, surrogate, steps, every, material=0):
    """steps steps (5 warm-up steps first), monitors read every `every`
    steps; returns (ms/step of the steps alone, monitor rows, health, active
    nodes).  The monitors are the reference's sequential sums over their
    particle lists (reportConsole order, so they match the oracle bit for
    bit), one thread per sum: on tens of millions of particles a read takes
    seconds, so it is kept out of the timed window, as the reference keeps
    its console output out of d_Runtime_Total."""
    import numpy as np
    from code_mpm_amd import Engine
    eng = Engine(device=0, order="pair")
    eng.load_scene(scene)
    if surrogate is not None:
        eng.set_surrogate(surrogate, material=material, residual_tolerance=1e-4)
    eng.step(scene.dt, 5)
    eng.synchronize()
    rows = []
    el = 0.0
    for _ in range(steps // every):
        t0 = time.perf_counter()
        eng.step(scene.dt, every)
        eng.synchronize()
        el += time.perf_counter() - t0
        m = eng.monitors()
        rows.append([m["position_y"], m["force_y"], m["strain_energy"], m["plastic_energy"], m["momentum_y"]])
    health = eng.health()
    act_mask = eng.get_node("ACTIVE")[0] != 0
    active = int(act_mask.sum())
    # tile order: the share of active nodes whose 8 cells lie in one
    # summation tile (finished by the tile kernel; the rest by the face pass)
    nn = np.asarray(scene.cells) + 1
    idx = np.nonzero(act_mask)[0]
    ix, iy, iz = idx % nn[0], (idx // nn[0]) % nn[1], idx // (nn[0] * nn[1])
    inner = float(np.mean((ix % 8 != 7) & (iy % 8 != 7) & (iz % 4 != 3))) if active else 0.0
    eng.close()
    return 1e3 * el / (every * (steps // every)), np.array(rows), health, active


def _curve_error(a, b):
    """max |a - b| / max |a| and ||a - b|| / ||a|| of two sampled curves."""
    import numpy as np
    peak, norm = float(np.abs(a).max()), float(np.linalg.norm(a))
    return {"max_over_peak": float(np.abs(a - b).max() / peak) if peak > 0 else None,
            "l2": float(np.linalg.norm(a - b) / norm) if norm > 0 else None}


def c4_bench(steps, fd_steps, every=100, fd_every=200, train_steps=1500):
    """C4: the half ring crushed by a foam impactor (scenes.half_ring_impact),
    exact FP64 return map against the learned one on the ring's steel.
    At the full ~48 M-particle size, `steps` steps each way: step time, the
    step's algorithmic HBM rate, the network's fallback rate and the error of
    the ring's energy curves.  The stress wave needs thousands of steps to
    reach the support, so the force-displacement comparison runs the same
    set-up at 4 h (about 0.8 M particles) for fd_steps steps: support reaction
    against platen displacement, exact vs learned.  The full-size runs' own
    support-reaction curve is reported beside it (force_displacement_full)."""
    import numpy as np
    from code_mpm_amd import scenes
    t = time.perf_counter()
    sc = scenes.half_ring_impact()
    print(f"# c4: {sc.n} particles, {sc.nnodes} nodes, setup {time.perf_counter() - t:.1f} s",
          file=sys.stderr, flush=True)
    s, hist = _train_surrogate(sc.materials[0], train_steps)
    full = {}
    for mode in ("exact", "surrogate"):
        ms, rows, health, active = _run_curve(sc, s if mode == "surrogate" else None, steps, every)
        full[mode] = dict(ms=ms, rows=rows, health=health, active=active)
        print(f"# c4 {mode}: {ms:.3f} ms/step", file=sys.stderr, flush=True)
    n, active = sc.n, full["exact"]["active"]
    step_bytes = 693 * n + 328 * active
    h = full["surrogate"]["health"]
    calls, falls = h["surrogate_calls"], h["surrogate_fallbacks"]
    e_rows, s_rows = full["exact"]["rows"], full["surrogate"]["rows"]
    y0_full = sc.fields["POSITION"][1][sc.dc_index].mean()
    del sc
    small = scenes.half_ring_impact(h=2.0e-4, depth_cells=23, foam_cells=20)
    fd = {}
    for mode in ("exact", "surrogate"):
        fd[mode] = _run_curve(small, s if mode == "surrogate" else None, fd_steps, fd_every)
    ea, sa = fd["exact"][1], fd["surrogate"][1]
    y0 = small.fields["POSITION"][1][small.dc_index].mean()
    return {
        "workload": f"half ring (steel, von Mises hardening) crushed by an aluminium foam impactor at "
                    f"{sc_speed(small)} m/s, {n} particles ({steps} steps each way, exact FP64 vs the learned "
                    "return map on the ring's steel, residual tolerance 1e-4 sigma_y)",
        "particles": n, "active_nodes": active, "steps": steps,
        "timing": "the steps alone, synchronised every monitor sample; monitor reads (sequential sums over "
                  "the lists) are outside it",
        "energy_samples_every_steps": every,
        "exact_ms_per_step": round(full["exact"]["ms"], 4), "surrogate_ms_per_step": round(full["surrogate"]["ms"], 4),
        "exact_particle_steps_per_s": round(n / (full["exact"]["ms"] * 1e-3), 1),
        "exact_algorithmic_gbs": round(step_bytes / (full["exact"]["ms"] * 1e-3) / 1e9, 1),
        "network_calls": calls, "fallback_rate": round(falls / max(calls, 1), 4),
        "plastic_points": full["exact"]["health"]["plastic_points"],
        "ring_strain_energy_error": _curve_error(e_rows[:, 2], s_rows[:, 2]),
        "ring_plastic_energy_error": _curve_error(e_rows[:, 3], s_rows[:, 3]),
        "force_displacement_full": {
            "what": f"support reaction against impactor displacement on the full-size runs, sampled every {every} "
                    "steps",
            "displacement_end_m": float(y0_full - e_rows[-1, 0]),
            "force_peak_exact_N": float(np.abs(e_rows[:, 1]).max()),
            "force_error": _curve_error(e_rows[:, 1], s_rows[:, 1])},
        "force_displacement": {
            "workload": f"the same set-up at h = 2e-4 m, {small.n} particles, {fd_steps} steps, "
                        f"support reaction sampled every {fd_every} steps",
            "displacement_end_m": float(y0 - ea[-1, 0]), "force_peak_exact_N": float(np.abs(ea[:, 1]).max()),
            "force_error": _curve_error(ea[:, 1], sa[:, 1]),
            "network_calls": fd["surrogate"][2]["surrogate_calls"],
            "fallback_rate": round(fd["surrogate"][2]["surrogate_fallbacks"] /
                                   max(fd["surrogate"][2]["surrogate_calls"], 1), 4)},
        "train_steps": train_steps, "train_loss_first_last": [float(hist[:10].mean()), float(hist[-10:].mean())],
    }


def beso_bench(n_designs):
    """One BESO iteration (Main.cpp:204-333) on the L-problem canvas of
    Main.cpp:127-145 (500 x 500 voxels of 0.1 mm, 57,636 particles per
    design, 16,000 steps to the end time), n_designs designs with different
    filter radii run as one batch of instances; then the objective update,
    the GPU smoothing filter and the soft-kill / restore pass per design."""
    from code_mpm_amd import beso, capi
    designs = [beso.Design(beso.l_problem(), radius=0.0012 * (1.0 + 0.25 * k)) for k in range(n_designs)]
    run = beso.BESO(designs, order="pair")
    t = time.perf_counter()
    lines = run.step()
    el = time.perf_counter() - t
    n = sum(sc.n for sc in run.last_scenes)
    steps = run.timing["steps"]
    cv = designs[0].canvas
    capi.eso_filter(cv.active, cv.eso, cv.objective, cv.nx, cv.ny, cv.d_Offset, designs[0].radius)
    t1 = time.perf_counter()
    for _ in range(5):
        capi.eso_filter(cv.active, cv.eso, cv.objective, cv.nx, cv.ny, cv.d_Offset, designs[0].radius)
    filt = (time.perf_counter() - t1) / 5
    out = {"designs": n_designs, "particles": n, "steps": steps, "iteration_s": round(el, 3),
           "simulate_s": round(run.timing["simulate_s"], 3),
           "value": round(n * steps / run.timing["simulate_s"], 1), "unit": "particle-steps/s",
           "filter_ms": round(1e3 * filt, 3), "filter_note": "mpm_eso_filter on 500 x 500 voxels incl. host copies",
           "removed": [d.removed for d in designs], "restored": [d.restored for d in designs],
           "log_first_design": lines[0].strip()}
    print(f"# beso: {n_designs} designs x {n // n_designs} particles, {steps} steps in {run.timing['simulate_s']:.2f} s, "
          f"filter {1e3 * filt:.2f} ms", file=sys.stderr, flush=True)
    return out, cv, designs[0].radius


def sc_speed(scene):
    return scene.info["impact_speed"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default node tile shape instead of the one mpm_autotune picks (one step)")
    ap.add_argument("--launch", default="eager", choices=["graph", "eager"],
                    help="time the K steps replayed from one HIP graph (recorded and run once first) or "
                         "launched kernel by kernel; the other is reported beside it")
    ap.add_argument("--no-late-window", dest="late_window", action="store_false",
                    help="skip the second timed window at steps 1951-2000 of the crush")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--order", default="pair", choices=["tile", "cell", "pair", "particle"],
                    help="node summation order (cell, pair, tile: fast; particle: the reference's one-thread order)")
    ap.add_argument("--variant-steps", type=int, default=10,
                    help="timed steps of the contact and single-pass loops on the same state (0: skip)")
    ap.add_argument("--batch-instances", type=int, default=256,
                    help="graded-foam batch size for the batched measurement (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=500,
                    help="graph-replayed steps of the batch (C2 names 500)")
    ap.add_argument("--surrogate-steps", type=int, default=2000,
                    help="steps of the seed-1 crush each way, exact and learned return map (0: skip)")
    ap.add_argument("--surrogate-seeds", default="2-9",
                    help="seeds of the crushes whose return-map samples train the network (a-b or a,b,...)")
    ap.add_argument("--surrogate-capture-steps", type=int, default=2000)
    ap.add_argument("--surrogate-rows", type=int, default=50_000_000,
                    help="captured rows over all training seeds")
    ap.add_argument("--surrogate-train-steps", type=int, default=20000)
    ap.add_argument("--surrogate-tol", type=float, default=1e-4,
                    help="C3: residual tolerance of the learned map's answers, in units of sigma_y")
    ap.add_argument("--c4-steps", type=int, default=1000,
                    help="C4 half-ring impact at ~48 M particles: steps each way, exact and learned (0: skip)")
    ap.add_argument("--c4-fd-steps", type=int, default=8000,
                    help="C4 force-displacement comparison at 4 h: steps each way")
    ap.add_argument("--beso-designs", type=int, default=4,
                    help="BESO designs batched in one iteration of the L-problem (0: skip)")
    ap.add_argument("--traffic-file", default=TRAFFIC_FILE,
                    help="rocprofv3 --pmc summary the roofline's measured traffic is read from")
    ap.add_argument("--copy-calibration-gib", type=float, default=0.0,
                    help="first run one device copy of this many GiB (the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes)")
    ap.add_argument("--cpu-baseline-child", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_child:
        _cpu_baseline_child(args.cpu_baseline_child)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding of this "
              "path is a design note, not a runnable configuration", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from code_mpm_amd import Engine, scenes

    if args.copy_calibration_gib > 0:
        nbytes = int(args.copy_calibration_gib * (1 << 30))
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print(f"# copy calibration: {nbytes} bytes read and written", file=sys.stderr, flush=True)

    t_setup = time.perf_counter()
    scene = scenes.honeycomb_crush()
    eng = Engine(device=0, order=args.order)
    eng.load_scene(scene)
    print(f"# scene {scene.name}: {scene.n} particles, setup {time.perf_counter() - t_setup:.1f} s",
          file=sys.stderr, flush=True)

    tile = eng.tile_shape()
    if not args.no_autotune:
        tile, _ = eng.autotune()   # one ordinary step that also picks the node tile shape
        print(f"# autotuned node tile {tile}", file=sys.stderr, flush=True)
    eng.step(scene.dt, args.warmup)
    eng.synchronize()
    torch.cuda.synchronize()

    def timed(launch):
        if launch == "graph":
            eng.step_graph(args.steps)  # records the K steps and runs them once
            eng.synchronize()
        torch.cuda.synchronize()
        t = time.perf_counter()
        if launch == "graph":
            eng.step_graph(args.steps)
        else:
            eng.step(scene.dt, args.steps)
        eng.synchronize()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    other = "eager" if args.launch == "graph" else "graph"
    elapsed_other = timed(other)
    elapsed = timed(args.launch)
    value = scene.n * args.steps / elapsed
    print(f"# timed {args.steps} steps in {elapsed:.4f} s ({args.launch}); {elapsed_other:.4f} s ({other})",
          file=sys.stderr, flush=True)

    kms = eng.step_timed_kernels(scene.dt, args.profile_steps)
    avg_ms = {k: v / args.profile_steps for k, v in kms.items()}
    act_mask = eng.get_node("ACTIVE")[0] != 0
    active = int(act_mask.sum())
    # tile order: the share of active nodes whose 8 cells lie in one
    # summation tile (finished by the tile kernel; the rest by the face pass)
    nn = np.asarray(scene.cells) + 1
    idx = np.nonzero(act_mask)[0]
    ix, iy, iz = idx % nn[0], (idx // nn[0]) % nn[1], idx // (nn[0] * nn[1])
    inner = float(np.mean((ix % 8 != 7) & (iy % 8 != 7) & (iz % 4 != 3))) if active else 0.0
    health = eng.health()
    # every step the engine has run so far (the autotune step, warm-up, both
    # launch modes' recordings and timed windows, the profiled steps): the
    # plastic-point counter covers all of them
    steps_total = eng.getTime_Cycle()
    plastic_frac = health["plastic_points"] / max(1, scene.n * steps_total)
    # a second timed window late in the 2000-step crush, when the walls have
    # started to yield: run on to step 1950, then time 50 steps
    late = None
    if args.late_window and steps_total < 1950:
        eng.step(scene.dt, 1950 - steps_total)
        eng.synchronize()
        p0, c0 = eng.health()["plastic_points"], eng.getTime_Cycle()
        t1 = time.perf_counter()
        eng.step(scene.dt, 50)
        eng.synchronize()
        el = time.perf_counter() - t1
        p1 = eng.health()["plastic_points"]
        late = {"first_step": c0 + 1, "steps": 50, "launch": "eager",
                "value": round(scene.n * 50 / el, 1), "unit": "particle-steps/s",
                "ms_per_step": round(1e3 * el / 50, 4),
                "plastic_fraction": round((p1 - p0) / (scene.n * 50), 5)}
        print(f"# late window (steps {c0 + 1}-{c0 + 50}): {late['ms_per_step']:.3f} ms/step, plastic fraction "
              f"{late['plastic_fraction']}", file=sys.stderr, flush=True)
    kernels = KERNELS[args.order]
    measured = read_traffic(args.traffic_file)
    per_kernel = {}
    for k, ms in avg_ms.items():
        if kernels[k] is None:
            continue
        per_p, per_n = PER_PARTICLE[k], PER_NODE[k]
        if k in FUSED_G2P:
            per_p, per_n = FUSED_G2P[k]
        if args.order == "tile" and k in ("p2g_partials", "remap_partials", "p2g_nodes", "remap_nodes"):
            node_k = "p2g_nodes" if k.startswith("p2g") else "remap_nodes"
            per_n = PER_NODE[node_k] * (inner if k.endswith("partials") else 1.0 - inner)
        b = int(round(per_p * scene.n + per_n * active + PER_PLASTIC.get(k, 0) * plastic_frac * scene.n))
        t = [measured.get(kn) for kn in kernels[k].split("+")]
        per_kernel[k] = dict(ms=round(ms, 4), kernel=kernels[k], algorithmic_bytes=b,
                             gbs=round(b / (ms * 1e-3) / 1e9, 1) if (b and ms > 0) else None,
                             traffic=round(sum(t)) if all(x is not None for x in t) else None)
    # the dominant kernel (one kernel, so the figure matches one line of the
    # rocprofv3 kernel summary)
    single = [k for k in per_kernel if "+" not in per_kernel[k]["kernel"]]
    dominant = max(single, key=avg_ms.get)
    dom_bytes = per_kernel[dominant]["algorithmic_bytes"]
    achieved = dom_bytes / (avg_ms[dominant] * 1e-3) / 1e9
    traffic = per_kernel[dominant]["traffic"]
    # the whole step's algorithmic bytes: the sum over its kernels
    step_bytes = sum(v["algorithmic_bytes"] for v in per_kernel.values())
    for k, v in per_kernel.items():
        print(f"# {k:13s} {v['ms']:8.3f} ms  {v['gbs']} GB/s  traffic {v['traffic']}", file=sys.stderr)

    variants = {}
    for var in (("doublepass_contact", "singlepass") if args.variant_steps > 0 else ()):
        eng.set_variant(var)
        eng.step(scene.dt, 2)
        eng.synchronize()
        t1 = time.perf_counter()
        eng.step(scene.dt, args.variant_steps)
        eng.synchronize()
        el = time.perf_counter() - t1
        variants[var] = {"value": round(scene.n * args.variant_steps / el, 1), "unit": "particle-steps/s",
                         "ms_per_step": round(1e3 * el / args.variant_steps, 4), "steps": args.variant_steps}
        print(f"# {var}: {variants[var]['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    eng.set_variant("doublepass")
    large = None
    if args.variant_steps > 0:
        # the large-strain grid-to-particle pass (logarithmic strains) on the same state
        eng.set_strain_formulation("large")
        eng.step(scene.dt, 2)
        eng.synchronize()
        t1 = time.perf_counter()
        eng.step(scene.dt, args.variant_steps)
        eng.synchronize()
        el = time.perf_counter() - t1
        kl = eng.step_timed_kernels(scene.dt, 3)
        large = {"value": round(scene.n * args.variant_steps / el, 1), "unit": "particle-steps/s",
                 "ms_per_step": round(1e3 * el / args.variant_steps, 4), "steps": args.variant_steps,
                 "kernel": "k_g2p_large<false>", "kernel_ms": round(kl["g2p_kinematics"] / 3, 4)}
        print(f"# large strain: {large['ms_per_step']:.3f} ms/step, pass {large['kernel_ms']:.3f} ms",
              file=sys.stderr, flush=True)
        eng.set_strain_formulation("small")
    eng.close()

    batch = batch_bench(args.batch_instances, args.batch_steps) if args.batch_instances > 0 else None
    sur = None
    if args.surrogate_steps > 0:
        sd = args.surrogate_seeds
        seeds = list(range(int(sd.split("-")[0]), int(sd.split("-")[1]) + 1)) if "-" in sd else \
            [int(v) for v in sd.split(",")]
        sur = surrogate_bench(args.surrogate_steps, seeds, args.surrogate_capture_steps, args.surrogate_rows,
                              args.surrogate_train_steps, tol=args.surrogate_tol)
    c4 = c4_bench(args.c4_steps, args.c4_fd_steps) if args.c4_steps > 0 else None
    beso_out, beso_canvas, beso_radius = beso_bench(args.beso_designs) if args.beso_designs > 0 else (None, None, None)

    cpu = None
    if not args.no_cpu_baseline:
        # every core this process may run on, up to the 16-core host share
        # of a one-GPU box
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        v_cpu, n_cpu, s_cpu, el_cpu, host = cpu_baseline(threads)
        cpu = {"value": round(v_cpu, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"x < 64 h slice of the same honeycomb ({n_cpu} particles, 64x512x32 cells), "
                         f"{s_cpu} steps after 3 warm-up, {el_cpu:.1f} s, OpenMP lock scatter, g++ -O2, "
                         f"one thread per core (OMP_PROC_BIND=close, OMP_PLACES=cores)",
               "host": host}
        print(f"# cpu baseline: {threads} threads on {host['affinity_cpus']} usable of {host['logical_cpus']} "
              f"logical CPUs: {v_cpu:.3e} particle-steps/s", file=sys.stderr, flush=True)
        if threads != 8:
            # the reference's own team size (PhysicsEngine.h:27), half the steps
            v8, n8, s8, el8, _ = cpu_baseline(8, steps=150)
            cpu["at_8_threads"] = {"value": round(v8, 1), "unit": "particle-steps/s", "cores": 8,
                                   "sample": f"same slice, {s8} steps, {el8:.1f} s"}
            print(f"# cpu baseline at 8 threads: {v8:.3e} particle-steps/s", file=sys.stderr, flush=True)
        if beso_out is not None:
            # the reference's all-pairs smoothing loop (Canvas2D_CC.cpp:302-341), one thread, same canvas
            from oracle import oracle as orc
            cv = beso_canvas
            t1 = time.perf_counter()
            orc.eso_filter(cv.active, cv.eso, cv.objective, cv.nx, cv.ny, cv.d_Offset, beso_radius)
            beso_out["cpu_filter_s"] = round(time.perf_counter() - t1, 3)
            print(f"# beso filter on the host (all pairs, 1 thread): {beso_out['cpu_filter_s']:.2f} s",
                  file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "J2 honeycomb crush, 512x512x32 cells, 8.0M particles, exact FP64 return map "
                               "(von Mises hardening), platen 10 m/s, seed 1; walls 5.75 h thick, not the 4 h "
                               "named (4 h gives 5.7M particles on this grid)",
                   "wall_thickness": "5.75 h (4 h gives 5.7 M particles on this grid; 5.75 h gives the named 8 M)",
                   "particles": scene.n, "active_nodes": active, "steps_in_plastic_fraction": steps_total,
                   "plastic_fraction": round(plastic_frac, 5),
                   "return_map": "exact", "summation_order": args.order,
                   "tile_interior_node_share": round(inner, 4) if args.order == "tile" else None,
                   "node_tile": list(tile) if tile is not None else None,
                   "launch": ("one HIP graph of the K steps, replayed" if args.launch == "graph"
                              else "kernel launches on one stream")},
        "other_launch": {"launch": other, "value": round(scene.n * args.steps / elapsed_other, 1),
                         "ms_per_step": round(1e3 * elapsed_other / args.steps, 4)},
        "roofline": {"bound": "hbm", "kernel": kernels[dominant], "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": (f"from {os.path.relpath(args.traffic_file, ROOT)} (rocprofv3 --pmc "
                                        "FETCH_SIZE and WRITE_SIZE, scaled by a copy of known size)")
                     if traffic is not None else None,
                     "algorithmic_bytes": dom_bytes,
                     "avg_launch_ms": round(avg_ms[dominant], 4)},
        "late_window": late,
        "step_effective_gbs": round(step_bytes / (elapsed / args.steps) / 1e9, 1),
        "kernels": per_kernel,
        "variants": variants,
        "large_strain": large,
        "batch": batch,
        "surrogate": sur,
        "c4": c4,
        "beso": beso_out,
        "cpu_baseline": cpu,
        "gpu_over_cpu": round(value / cpu["value"], 1) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
