"""Benchmark: particle-updates/s of the explicit MPM step on one MI355X.

Workload: the 3D column collapse at 15,990,784 points (256x256x128 grid over 2x2x1 m,
h = 7.8125 mm, 2^3 points per cell jittered +-0.25 cell, E = 1e7 Pa, nu = 0.3,
rho = 1500, dt = 1e-5, gravity, CartesianBox walls), synthetic and seeded.

A step is one full ExplicitUSL step over all points with the state resident in HBM:
sort (every 16 steps), map + P2G, grid update, G2P + stress. Both forms run on the same
input: the plain FP64 form and the precision form (FP32 particle state and stress, FP64
nodal accumulation, grid update and contact normals). `value` is the precision form's
active points x timed steps / wall time of the timed steps; `forms` holds both and
`fp32_vs_fp64` their ratio.

The JSON line also carries
  roofline      the dominant kernel's algorithmic HBM bytes per launch / its average launch
                time (HIP events on the library's stream), against 8 TB/s
  cpu_baseline  the oracle's threaded twin of ThreadPoolExplicitUSL (oracle/, -O3
                -march=native) on a bounded sample of the same workload, on this host
  collision     the two-block collision on 256^3 (7.3M points, two bodies, ContactHuang),
                plain FP64 form, as configured and with contact engaged
  batched       512 independent 2D block drops (16^2..256^2 grids) in one handle, precision
                form, per-simulation dt, HIP-graph sort periods, autotuned launch shapes:
                solves/s and particle-updates/s against the same simulations one at a time
  served        128 3D block drops / collisions admitted into 36 slots of three capacity
                classes (each request into the smallest free slot that fits), continuous
                batching: capacity (all queued at once) against the same requests one at a
                time (sampled), p50/p99 latency under Poisson arrivals at 75% of that
                capacity, and the mix's bound on the speed-up
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s, 16M-pt 3D (FP32 vs FP64); batched solves/s + p99 latency"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per launch, 3D, as (bytes per active point, bytes per active node): every
# plane the kernel must read once and every plane it must write once
BYTES = {
    "fp64": {
        "pass": (224, 104),     # read x3 v3 T6 V m extent, write x3 v3 T6 V; read v3 dv3, write m p3 f3
        "grid": (0, 104),       # read m p3 f3, write v3 dv3
        "p2g": (120, 56),       # read x3 v3 T6 V m extent; write m p3 f3
        "g2p": (216, 48),       # read x3 v3 T6 V extent, write x3 v3 T6 V; read v3 dv3
        "step": (224, 208),     # one particle pass + one grid update
    },
    "fp32": {
        "grid": (0, 80),        # read m p3 f3 (FP64), write v3 dv3 (FP32)
        "p2g": (68, 56),        # read cell q3 v3 T6 lnJ V0 m extent (4 B each); write m p3 f3 (FP64)
        "g2p": (116, 24),       # read cell q3 v3 T6 lnJ extent, write cell q3 v3 T6 lnJ; read v3 dv3 (FP32)
        "step": (124, 160),     # each plane read once and the updated ones written once
    },
}
CLASSES = {0: "pass", 1: "grid", 2: "sort", 3: "p2g", 4: "g2p"}


def host_cpu():
    """CPU model (/proc/cpuinfo), logical CPUs, and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = nproc
    return model, nproc, usable


def _time_steps(ora, dt, threads, min_s, min_steps):
    """Steps of the oracle until at least `min_s` seconds and `min_steps` steps have passed."""
    steps = 0
    t0 = time.perf_counter()
    while True:
        ora.step(dt, 1, nthreads=threads)
        steps += 1
        el = time.perf_counter() - t0
        if steps >= min_steps and el >= min_s:
            return steps, el


def cpu_baseline(threads):
    """The oracle's threaded twin of ThreadPoolExplicitUSL (persistent worker pool, -O3
    -march=native) and its serial ExplicitUSL on the 1/16-footprint sample of the same
    column (same h, E, dt), each timed for a few seconds on this host."""
    from mpm_3d_amd import workloads as W
    from tests._harness import oracle_of
    model, nproc, usable = host_cpu()
    w = W.column_3d_sample()
    ora = oracle_of(w, native=True)
    ora.step(w.dt, 1, nthreads=threads)  # warm-up: allocations, worker start
    st, el = _time_steps(ora, w.dt, threads, 4.0, 2)
    st1, el1 = _time_steps(ora, w.dt, 1, 2.0, 1)
    return {
        "value": w.n_points * st / el,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{w.n_points} points, 64x64x128 grid (same h, material, dt as the 16M column): "
                  f"{st} steps of the threaded ExplicitUSL restatement on {threads} threads in {el:.1f} s, "
                  f"{st1} serial step(s) in {el1:.1f} s",
        "serial_value": w.n_points * st1 / el1,
        "host": {"cpu_model": model, "nproc": nproc, "usable_cpus": usable,
                 "threads_used": threads},
    }


def run_form(w, form, args, settle=None, traffic_key="column_3d_16M"):
    """Time `args.steps` steps of one form; returns its summary."""
    from mpm_3d_amd.gpu import (MPM_OPT_AUTOTUNE, MPM_OPT_GRID_WGS, MPM_OPT_P2G_GRID, MPM_OPT_PASS_GRID,
                                MPM_OPT_SORT_INTERVAL, MPM_PREC_FP32, MPM_PREC_FP64, GpuSim)
    prec = MPM_PREC_FP32 if form == "fp32" else MPM_PREC_FP64
    sim = GpuSim(w, precision=prec, sort_interval=args.sort_interval, fused=args.fused)
    # the column's first steps are faster than later ones (points spread over more dual cells
    # as it deforms); `settle` untimed steps start the timed window in the steadier regime.
    # The autotuner picks the launch shapes and the sort interval on the first of them.
    settle = args.settle if settle is None else settle
    if not args.no_autotune and settle + args.warmup >= 90:  # the tuning needs 84 steps of its own
        sim.set_option(MPM_OPT_AUTOTUNE, 1)
    if settle + args.warmup:
        sim.step(w.dt, settle + args.warmup)
    tuned = {"sort_interval": sim.get_option(MPM_OPT_SORT_INTERVAL), "pass_grid": sim.get_option(MPM_OPT_PASS_GRID),
             "p2g_grid": sim.get_option(MPM_OPT_P2G_GRID), "grid_wgs": sim.get_option(MPM_OPT_GRID_WGS),
             "autotuned": sim.get_option(MPM_OPT_AUTOTUNE) == 2}
    sim.synchronize()
    sim.set_timing(True)
    t0 = time.perf_counter()
    sim.step(w.dt, args.steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    active_nodes, active_points = sim.stats()
    kern = {}
    for cls, name in CLASSES.items():
        n, ms = sim.timing(cls)
        if n:
            kern[name] = {"launches": n, "avg_ms": ms / n, "total_ms": ms}
    sim.close()
    print(f"[bench] {form}: {args.steps} steps in {wall:.3f} s; active points {active_points}, "
          f"nodes {active_nodes}", file=sys.stderr, flush=True)
    table = BYTES[form]
    dom = max((k for k in kern if k in table), key=lambda k: kern[k]["total_ms"])
    bp, bn = table[dom]
    bytes_per_launch = bp * active_points + bn * active_nodes
    achieved = bytes_per_launch / (kern[dom]["avg_ms"] * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_{form}_{traffic_key}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    step_bytes = (table["step"][0] * active_points + table["step"][1] * active_nodes) * args.steps
    return {
        "value": active_points * args.steps / wall,
        "ms_per_step": wall / args.steps * 1e3,
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "step_hbm_fraction": round(step_bytes / wall / 1e9 / HBM_PEAK_GBS, 4),
        "kernels": kern,
        "tuned": tuned,
    }


def run_collision(args):
    """Two-block collision on 256^3 (two 0.3 m cubes at 2^3 points per cell, 7.2M points,
    +-1 m/s, E = 1e6, ContactHuang mu = 0.4, CartesianBox walls, dt = 2e-5) in the plain
    FP64 form, as configured (centres 0.4 m apart: the cubes do not meet within its 1000
    steps) and with the faces one cell apart so that contact is engaged from the first
    steps; `args.steps` timed steps after `args.warmup`."""
    from mpm_3d_amd import workloads as W
    out = {}
    for name, gap in (("as_configured", None), ("contact_engaged", 1)):
        w = W.two_block_3d(n_cells=256, gap_cells=gap)
        r = run_form(w, "fp64", args, settle=0, traffic_key="collision_3d_256")
        out[name] = {"workload": f"{w.name}: {w.n_points} points, centres "
                                 f"{'0.4 m apart' if gap is None else 'one cell apart at the faces'}",
                     "value": r["value"], "ms_per_step": r["ms_per_step"], "roofline": r["roofline"],
                     "kernels": r["kernels"]}
        print(f"[bench] collision {name}: {r['value']:.3e} particle-updates/s, {r['ms_per_step']:.3f} ms/step",
              file=sys.stderr, flush=True)
    return out


def run_batched(args):
    """Batched 2D ensemble: 512 block drops (seeded) in one handle, precision form,
    per-simulation dt, graph-replayed sort periods, autotuned launch shapes, `ens_steps`
    steps each; against the same simulations run one at a time (timed on a sample)."""
    from mpm_3d_amd import workloads as W
    from mpm_3d_amd.gpu import MPM_OPT_AUTOTUNE, MPM_OPT_GRAPH, MPM_OPT_GRAPH_CAPTURES, MPM_OPT_GRID_WGS, \
        MPM_OPT_P2G_GRID, MPM_OPT_PASS_GRID, MPM_OPT_SORT_INTERVAL, MPM_PREC_FP32, GpuSim
    ws = W.ensemble_2d(n_sims=512, steps=args.ens_steps)
    npts = sum(w.n_points for w in ws)
    # Both legs are timed the same way: handle creation and upload (setup) apart from the
    # stepping, each reported.
    ts = time.perf_counter()
    sim = GpuSim(ws, precision=MPM_PREC_FP32)
    sim.set_option(MPM_OPT_GRAPH, 0 if args.no_graph else 1)
    sim.set_option(MPM_OPT_AUTOTUNE, 1)
    sim.synchronize()
    setup = time.perf_counter() - ts
    t0 = time.perf_counter()
    sim.step_batch([w.dt for w in ws], args.ens_steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    tuned = {"pass_grid": sim.get_option(MPM_OPT_PASS_GRID), "p2g_grid": sim.get_option(MPM_OPT_P2G_GRID),
             "grid_wgs": sim.get_option(MPM_OPT_GRID_WGS), "sort_interval": sim.get_option(MPM_OPT_SORT_INTERVAL),
             "graph_captures": sim.get_option(MPM_OPT_GRAPH_CAPTURES)}
    sim.close()
    print(f"[bench] batched: {len(ws)} sims, {npts} points, {args.ens_steps} steps in {wall:.2f} s "
          f"(setup {setup:.2f} s)", file=sys.stderr, flush=True)
    # one at a time: every 32nd simulation on its own handle (same form, same steps)
    sample = ws[::32]
    alone = alone_setup = 0.0
    for w in sample:
        ts = time.perf_counter()
        one = GpuSim(w, precision=MPM_PREC_FP32)
        one.synchronize()
        t1 = time.perf_counter()
        one.step(w.dt, args.ens_steps)
        one.synchronize()
        alone += time.perf_counter() - t1
        alone_setup += t1 - ts
        one.close()
    alone_pts = sum(w.n_points for w in sample)
    out = {
        "workload": f"ensemble_2d: 512 block drops, grid 16^2..256^2, {npts} points total, {args.ens_steps} steps each",
        "solves_per_s": len(ws) / wall,
        "particle_updates_per_s": npts * args.ens_steps / wall,
        "wall_s": wall,
        "setup_s": setup,
        "tuned": tuned,
        "one_at_a_time": {"sample": len(sample), "solves_per_s": len(sample) / alone,
                          "particle_updates_per_s": alone_pts * args.ens_steps / alone,
                          "stepping_s": alone, "setup_s": alone_setup,
                          "note": "each sampled simulation on its own handle; creation and upload timed "
                                  "apart from the stepping, as in the batched leg"},
    }
    out["speedup_solves"] = out["solves_per_s"] / out["one_at_a_time"]["solves_per_s"]
    out["speedup_particle_updates"] = out["particle_updates_per_s"] / out["one_at_a_time"]["particle_updates_per_s"]
    # the same ratio with setup counted on both sides (the sample scaled to 512 simulations)
    scale = len(ws) / len(sample)
    out["speedup_solves_with_setup"] = ((alone + alone_setup) * scale) / (wall + setup)
    return out


# Slot classes of the served leg: the largest request each N range of the served-stream
# generator draws (drops of up to min(N/2 - 2, 62) cells a side, collisions of two cubes of up
# to min((N - 6)/2 - 1, 49) cells, 8 points per cell), body 0 holding either, body 1 a cube
SERVE_CLASSES = ((64, (216_000, 175_616)), (96, (778_688, 681_472)), (128, (1_906_624, 941_192)))


def run_served(args, full_rate=None):
    """Served 3D stream: `serve_n` requests (seeded block drops / two-block collisions,
    N^3 grids 32..128, 1e4..2e6 points, 500..5000 steps) through a pool of slot classes
    (N_cap 64, 96 and 128, `serve_slots` slots each, fewer slots than requests) with
    continuous batching: each request takes the smallest free slot that fits it, and waits
    when none does. Capacity: all requests queued at t = 0. Latency: Poisson arrivals at
    `serve_load` x the measured capacity (solves/s); latency = completion - arrival, the
    result downloaded. One at a time: every `serve_sample`-th request on a handle of its
    own (create, upload, step, download), its time scaled to the whole stream. The bound:
    one-at-a-time time over the time the stream's particle-updates take at `full_rate`
    (the 16M column's rate, the GPU full), which no schedule can beat."""
    from mpm_3d_amd import serve as S
    from mpm_3d_amd import workloads as W
    from mpm_3d_amd.gpu import MPM_OPT_SORT_INTERVAL, MPM_PREC_FP32
    reqs = W.serve_requests(args.serve_n)
    counts = [int(c) for c in args.serve_slots.split(",")]
    classes = [S.SlotClass(c, N, caps) for c, (N, caps) in zip(counts, SERVE_CLASSES)]
    pu = sum(w.n_points * st for w, st in reqs)
    pool = S.SlotPool(classes=classes, dim=3, precision=MPM_PREC_FP32, graphs=not args.no_graph,
                      autotune=args.serve_autotune)
    sat = S.run_stream(pool, reqs, [0.0] * len(reqs), chunk_max=64)
    cap_solves = len(reqs) / sat["makespan"]
    rate = args.serve_load * cap_solves
    arrivals = S.poisson_arrivals(len(reqs), rate)
    res = S.run_stream(pool, reqs, arrivals, chunk_max=64)
    footprint = sum(pool.footprint)
    sort_interval = pool.sim.get_option(MPM_OPT_SORT_INTERVAL)
    pool.close()
    lat = np.array([f - a for f, a in zip(res["finish"], res["arrival"])])
    print(f"[bench] served: {len(reqs)} requests on {pool.n_slots} slots ({footprint / 1e9:.1f} GB); capacity run "
          f"{sat['makespan']:.2f} s; at {rate:.2f}/s p50 {np.percentile(lat, 50):.3f} s p99 "
          f"{np.percentile(lat, 99):.3f} s", file=sys.stderr, flush=True)
    sample = list(range(0, len(reqs), args.serve_sample))
    t1 = time.perf_counter()
    for i in sample:
        S.run_alone(*reqs[i])
    alone_sample = time.perf_counter() - t1
    alone = alone_sample * len(reqs) / len(sample)
    print(f"[bench] served: one at a time {alone_sample:.2f} s for {len(sample)} requests", file=sys.stderr,
          flush=True)
    out = {
        "workload": f"{len(reqs)} requests (block drop / two-block collision, N^3 = 32^3..128^3, "
                    f"{min(w.n_points for w, _ in reqs)}..{max(w.n_points for w, _ in reqs)} points, "
                    f"500..5000 steps, {pu:.3e} particle-updates in all)",
        "slots": {"classes": [{"count": c.count, "N_cap": c.N_cap, "body_caps": list(c.body_caps)} for c in classes],
                  "total": pool.n_slots, "footprint_bytes": footprint, "sort_interval": sort_interval},
        "statuses": {st: sat["status"].count(st) for st in set(sat["status"])},
        "capacity": {"solves_per_s": cap_solves, "particle_updates_per_s": pu / sat["makespan"],
                     "makespan_s": sat["makespan"], "stepping_s": sat["step_s"], "admission_s": sat["admin_s"],
                     "admission_parts_s": {"load": sat["load_s"], "result": sat["result_s"], "free": sat["free_s"]},
                     "host_work_overlapped_s": sat["host_overlapped_s"], "overlap": sat["overlap"],
                     "chunks": sat["chunks"]},
        "latency": {"arrival_rate_per_s": rate, "load": args.serve_load,
                    "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                    "solves_per_s": len(reqs) / res["makespan"], "makespan_s": res["makespan"]},
        "one_at_a_time": {"solves_per_s": len(reqs) / alone, "particle_updates_per_s": pu / alone,
                          "wall_s_estimated": alone, "sample": len(sample), "sample_wall_s": alone_sample},
    }
    out["speedup_solves"] = cap_solves / out["one_at_a_time"]["solves_per_s"]
    if full_rate:
        t_min = pu / full_rate
        out["bound"] = {"full_gpu_particle_updates_per_s": full_rate, "min_makespan_s": t_min,
                        "max_speedup": alone / t_min}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle", type=int, default=200,
                    help="untimed steps of the column before the warm-up, so the timed window is not its start")
    ap.add_argument("--precision", choices=["both", "fp32", "fp64"], default="both")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--sort-interval", type=int, default=16)
    ap.add_argument("--fused", action="store_true", help="one fused G2P+P2G pass per step instead of two passes")
    ap.add_argument("--no-autotune", action="store_true",
                    help="column and collision legs with the default launch shapes and sort interval")
    ap.add_argument("--legs", default="column,collision,batched,served",
                    help="comma list of: column (16M column, both forms), collision (256^3 two-block "
                         "collision, FP64), batched (512-sim 2D ensemble), served")
    ap.add_argument("--ens-steps", type=int, default=2000)
    ap.add_argument("--no-graph", action="store_true",
                    help="batched and served legs without graph replay (used for every profiled run)")
    ap.add_argument("--serve-n", type=int, default=128)
    ap.add_argument("--serve-slots", default="16,12,8",
                    help="slots of the N_cap 64, 96 and 128 classes of the served pool")
    ap.add_argument("--serve-sample", type=int, default=4,
                    help="time every n-th served request alone for the one-at-a-time figure")
    ap.add_argument("--serve-load", type=float, default=0.75, help="arrival rate as a fraction of capacity")
    ap.add_argument("--serve-autotune", action="store_true", help="served pool tunes its launch shapes")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the plain copy kernel (1 GiB, 10 launches) used to calibrate counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path is "
              "design-only in this build", file=sys.stderr)
        return 2

    from mpm_3d_amd import workloads as W
    from mpm_3d_amd.gpu import copy_probe

    if args.copy_probe:
        print(json.dumps({"copy_probe_gbps": copy_probe(1 << 30, 10), "bytes_per_launch": 2 * (1 << 30)}))
        return 0
    hbm_copy = copy_probe(1 << 30, 20)
    w = W.column_3d(wide=True)
    print(f"[bench] workload {w.name}: {w.n_points} points, grid {w.N}", file=sys.stderr, flush=True)
    forms = ["fp64", "fp32"] if args.precision == "both" else [args.precision]
    res = {f: run_form(w, f, args) for f in forms}
    main_form = "fp32" if "fp32" in res else "fp64"
    r = res[main_form]
    out = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 accumulation" if main_form == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": w.name, "points": int(w.n_points), "grid": list(w.N), "dt": w.dt,
                   "timed_window": f"steps {args.settle + args.warmup + 1}..{args.settle + args.warmup + args.steps}",
                   "form": "precision (FP32 state, FP64 accumulation)" if main_form == "fp32" else "plain FP64",
                   "sort_interval": r["tuned"]["sort_interval"], "autotuned": r["tuned"]["autotuned"],
                   "particle_pass": "fused" if args.fused else "split"},
        "roofline": r["roofline"],
        "step_hbm_fraction": r["step_hbm_fraction"],
        "hbm_copy_gbps": round(hbm_copy, 1),
        "kernels": r["kernels"],
        "forms": res,
    }
    if len(res) == 2:
        out["fp32_vs_fp64"] = round(res["fp32"]["value"] / res["fp64"]["value"], 3)
    legs = set(args.legs.split(","))
    if "collision" in legs:
        out["collision"] = run_collision(args)
    if "batched" in legs:
        out["batched"] = run_batched(args)
    if "served" in legs:
        out["served"] = run_served(args, res["fp32"]["value"] if "fp32" in res else None)
    if not args.no_cpu:
        # one GPU's share of the host: at most 16 CPUs, fewer if the process may use fewer
        threads = args.cpu_threads or min(16, host_cpu()[2])
        print(f"[bench] cpu baseline with {threads} threads ({host_cpu()})", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
