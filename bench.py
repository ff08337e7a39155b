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

# Algorithmic bytes (FP64 words counted once per pass that needs them):
#   g2p_strain: reads x3 F9 sigma6 eps6 epsp6 R W4 V0 (36 words) and writes
#     x3 F9 V sigma6 eps6 epsp6 R W4 (36 words), plus material id (4 B) and
#     control flags (1 B); per active node it reads u3 v3 m (7 words).
#   p2g_grid: reads x3 v3 m V sigma6 (14 words, +3 with external force);
#     per active node writes m p3 f3 (7 words, the P2G flush) and reads
#     fixed/friction and writes p3 f3 f/m3 reaction (13 words, the grid update).
#   g2p_velocity: reads x3 v3, writes v3; per active node reads f/m3 m.
#   remap_grid: reads x3 v3 m (7 words); per active node reads m p3, writes v3 p/m3.
PER_PARTICLE = {"g2p_strain": 72 * 8 + 5, "p2g_grid": 14 * 8, "g2p_velocity": 9 * 8, "remap_grid": 7 * 8,
                "bin": 0}
PER_NODE = {"g2p_strain": 7 * 8, "p2g_grid": 20 * 8, "g2p_velocity": 4 * 8, "remap_grid": 10 * 8, "bin": 0}
# kernels of each phase, per summation order (names as rocprofv3 reports them)
KERNELS = {
    "cell": {"bin": ["k_scan_reduce", "k_scan_blocks", "k_scan_apply", "k_scatter", "k_cell_sort"],
             "p2g_grid": ["k_cell_partials<false>", "k_node_cells<false>"], "g2p_velocity": ["k_g2p_velocity"],
             "remap_grid": ["k_cell_partials<true>", "k_node_cells<true>"], "g2p_strain": ["k_g2p_strain<0, 0>"]},
    "particle": {"bin": ["k_scan_reduce", "k_scan_blocks", "k_scan_apply", "k_scatter", "k_cell_sort"],
                 "p2g_grid": ["k_p2g_grid"], "g2p_velocity": ["k_g2p_velocity"], "remap_grid": ["k_remap_grid"],
                 "g2p_strain": ["k_g2p_strain<0, 0>"]},
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_honeycomb8m.txt")


def read_traffic():
    """HBM bytes per launch per kernel, from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE scaled by a copy of known size)."""
    out = {}
    if not os.path.exists(TRAFFIC_FILE):
        return out
    for line in open(TRAFFIC_FILE):
        if line.startswith("#") or "hbm_bytes_per_launch[" not in line:
            continue
        tail = line.split("hbm_bytes_per_launch[", 1)[1]
        name, val = tail.split("]:", 1)
        out[name.strip()] = float(val)
    return out


def cpu_baseline(scene, threads, steps=20, warmup=3):
    from code_mpm_amd import scenes
    from oracle.oracle import OracleEngine
    sample = scenes.crop_x(scene, 64)
    o = OracleEngine(threads=threads)
    o.load_scene(sample)
    o.step(sample.dt, warmup)
    t0 = time.perf_counter()
    o.step(sample.dt, steps)
    el = time.perf_counter() - t0
    o.close()
    return sample.n * steps / el, sample.n, steps, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--order", default="cell", choices=["cell", "particle"],
                    help="node summation order (cell: fast, particle: the reference's one-thread order)")
    ap.add_argument("--variant-steps", type=int, default=10,
                    help="timed steps of the contact and single-pass loops on the same state (0: skip)")
    ap.add_argument("--copy-calibration-gib", type=float, default=0.0,
                    help="first run one device copy of this many GiB (the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
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

    eng.step(scene.dt, args.warmup)
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(scene.dt, args.steps)
    eng.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = scene.n * args.steps / elapsed
    print(f"# timed {args.steps} steps in {elapsed:.4f} s", file=sys.stderr, flush=True)

    phases = eng.step_timed(scene.dt, args.profile_steps)
    avg_ms = {k: v / args.profile_steps for k, v in phases.items()}
    active = int(eng.get_node("ACTIVE").sum())
    health = eng.health()
    steps_total = args.warmup + args.steps + args.profile_steps
    plastic_frac = health["plastic_points"] / max(1, scene.n * steps_total)
    kernels = KERNELS[args.order]
    measured = read_traffic() if args.order == "cell" else {}
    per_kernel = {}
    for k, ms in avg_ms.items():
        b = PER_PARTICLE[k] * scene.n + PER_NODE[k] * active
        t = [measured.get(kn) for kn in kernels[k]]
        per_kernel[k] = dict(ms=round(ms, 4), kernels=kernels[k], algorithmic_bytes=b,
                             gbs=round(b / (ms * 1e-3) / 1e9, 1) if b else None,
                             traffic=round(sum(t)) if t and all(x is not None for x in t) else None)
    # the dominant single kernel (a phase made of one kernel), so the figure
    # matches one line of the rocprofv3 kernel summary
    single = [k for k in avg_ms if len(kernels[k]) == 1 and PER_PARTICLE[k]]
    dominant = max(single, key=avg_ms.get)
    dom_bytes = per_kernel[dominant]["algorithmic_bytes"]
    achieved = dom_bytes / (avg_ms[dominant] * 1e-3) / 1e9
    traffic = per_kernel[dominant]["traffic"]
    step_bytes = 768 * scene.n + 272 * active
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

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        v_cpu, n_cpu, s_cpu, el_cpu = cpu_baseline(scene, threads)
        cpu = {"value": round(v_cpu, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"x < 64 h slice of the same honeycomb ({n_cpu} particles, 64x512x32 cells), "
                         f"{s_cpu} steps after 3 warm-up, {el_cpu:.1f} s, OpenMP lock scatter, g++ -O2"}

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "J2 honeycomb crush, 512x512x32 cells, 8.0M particles, exact FP64 return map "
                               "(von Mises hardening), platen 10 m/s, seed 1",
                   "particles": scene.n, "active_nodes": active, "plastic_fraction": round(plastic_frac, 5),
                   "return_map": "exact", "summation_order": args.order},
        "roofline": {"bound": "hbm", "kernel": kernels[dominant][0], "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes": dom_bytes,
                     "avg_launch_ms": round(avg_ms[dominant], 4)},
        "step_effective_gbs": round(step_bytes / (elapsed / args.steps) / 1e9, 1),
        "phases": per_kernel,
        "variants": variants,
        "cpu_baseline": cpu,
        "gpu_over_cpu": round(value / cpu["value"], 1) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
